// Mat.hpp -- the small row-major double matrix used by the C++ mirror of the
// reference classes.  Its memory is exactly that of a continuous CV_64F
// cv::Mat (rows x cols doubles, row-major), so a caller holding cv::Mat data
// can wrap it with Mat(rows, cols, ptr) without conversion.
#pragma once

#include <cstddef>
#include <stdexcept>
#include <vector>

namespace mkfbt {

class Mat {
public:
    int rows = 0, cols = 0;
    std::vector<double> data;

    Mat() = default;
    Mat(int r, int c, double v = 0.0) : rows(r), cols(c), data((size_t)r * c, v) {}
    Mat(int r, int c, const double *p) : rows(r), cols(c), data(p, p + (size_t)r * c) {}

    static Mat zeros(int r, int c) { return Mat(r, c, 0.0); }
    static Mat eye(int n) {
        Mat m(n, n, 0.0);
        for (int i = 0; i < n; ++i) m(i, i) = 1.0;
        return m;
    }

    double &operator()(int r, int c) { return data[(size_t)r * cols + c]; }
    double operator()(int r, int c) const { return data[(size_t)r * cols + c]; }
    double &at(int r, int c) { check(r, c); return (*this)(r, c); }
    double at(int r, int c) const { check(r, c); return (*this)(r, c); }
    double *ptr() { return data.data(); }
    const double *ptr() const { return data.data(); }
    bool empty() const { return data.empty(); }

    Mat t() const {
        Mat o(cols, rows);
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < cols; ++c) o(c, r) = (*this)(r, c);
        return o;
    }
    Mat row(int r) const { return Mat(1, cols, data.data() + (size_t)r * cols); }
    Mat col(int c) const {
        Mat o(rows, 1);
        for (int r = 0; r < rows; ++r) o(r, 0) = (*this)(r, c);
        return o;
    }
    Mat rowRange(int r0, int r1) const { return Mat(r1 - r0, cols, data.data() + (size_t)r0 * cols); }

private:
    void check(int r, int c) const {
        if (r < 0 || r >= rows || c < 0 || c >= cols) throw std::out_of_range("Mat::at index out of range");
    }
};

// Raised for every negative status of the C ABI (the reference's OpenCV calls
// throw cv::Exception on argument errors).
class PfException : public std::runtime_error {
public:
    PfException(int code, const std::string &msg) : std::runtime_error(msg), code(code) {}
    int code;
};

}  // namespace mkfbt
