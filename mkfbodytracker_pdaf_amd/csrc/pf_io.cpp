// pf_io.cpp -- host input formats and host RNG of the particle-filter library:
// the OpenCV FileStorage YAML matrices the reference reads with
// cv::FileStorage (src/pfPose.cpp:34-53), the ROS camera_info YAML (cal.yml),
// and the Philox4x32-10 stream that supplies the host uniforms.
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/pf_rbpf.h"
#include "pf_internal.hpp"

namespace mkfbt {

void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void philox_u32(uint64_t seed, uint32_t stream, uint64_t first, int64_t count, uint32_t *out)
{
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint64_t cached = ~0ull;
    uint32_t o[4] = {0, 0, 0, 0};
    for (int64_t i = 0; i < count; ++i) {
        const uint64_t w = first + (uint64_t)i;
        const uint64_t blk = w >> 2;
        if (blk != cached) {
            const uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), stream, 0u};
            philox4x32_10(ctr, key, o);
            cached = blk;
        }
        out[i] = o[w & 3];
    }
}

namespace {

std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

int read_lines(const std::string &path, std::vector<std::string> &lines)
{
    std::ifstream f(path);
    if (!f) return set_error(PF_ERR_IO, "cannot open %s", path.c_str());
    std::string l;
    while (std::getline(f, l)) lines.push_back(l);
    return PF_OK;
}

// Index of the top-level line "key:" (no indentation), or -1.
int find_key(const std::vector<std::string> &lines, const std::string &key)
{
    for (size_t i = 0; i < lines.size(); ++i) {
        const std::string &l = lines[i];
        if (l.empty() || std::isspace((unsigned char)l[0])) continue;
        const size_t c = l.find(':');
        if (c == std::string::npos) continue;
        if (trim(l.substr(0, c)) == key) return (int)i;
    }
    return -1;
}

}  // namespace

int yml_read_matrix(const std::string &path, const std::string &key, YmlMatrix &out)
{
    std::vector<std::string> lines;
    int st = read_lines(path, lines);
    if (st) return st;
    const int k = find_key(lines, key);
    if (k < 0) return set_error(PF_ERR_IO, "key '%s' not found in %s", key.c_str(), path.c_str());
    out = YmlMatrix();
    bool have_rows = false, have_cols = false;
    std::string data;
    bool in_data = false;
    for (size_t i = (size_t)k + 1; i < lines.size(); ++i) {
        const std::string &l = lines[i];
        if (!l.empty() && !std::isspace((unsigned char)l[0])) break;  // next top-level key
        std::string t = trim(l);
        if (in_data) {
            data += " " + t;
            if (t.find(']') != std::string::npos) break;
            continue;
        }
        const size_t c = t.find(':');
        if (c == std::string::npos) continue;
        const std::string name = trim(t.substr(0, c));
        const std::string val = trim(t.substr(c + 1));
        if (name == "rows") { out.rows = std::atoi(val.c_str()); have_rows = true; }
        else if (name == "cols") { out.cols = std::atoi(val.c_str()); have_cols = true; }
        else if (name == "dt") { out.dt = val.empty() ? 'd' : val[0]; }
        else if (name == "data") {
            data = val;
            if (val.find(']') == std::string::npos) in_data = true;
            else break;
        }
    }
    if (!have_rows || !have_cols) return set_error(PF_ERR_IO, "'%s' in %s is not a matrix node", key.c_str(), path.c_str());
    const size_t a = data.find('['), b = data.rfind(']');
    if (a == std::string::npos || b == std::string::npos || b < a)
        return set_error(PF_ERR_IO, "'%s' in %s has no data list", key.c_str(), path.c_str());
    std::string body = data.substr(a + 1, b - a - 1);
    for (char &ch : body)
        if (ch == ',') ch = ' ';
    std::istringstream is(body);
    std::string tok;
    while (is >> tok) {
        char *end = nullptr;
        const double v = std::strtod(tok.c_str(), &end);
        if (end == tok.c_str()) return set_error(PF_ERR_IO, "bad number '%s' in '%s'", tok.c_str(), key.c_str());
        // dt f: the values are FP32 in the file; round them as cv::Mat CV_32F does.
        out.data.push_back(out.dt == 'f' ? (double)(float)v : v);
    }
    if ((long long)out.data.size() != (long long)out.rows * out.cols)
        return set_error(PF_ERR_IO, "'%s' in %s: %zu values for %d x %d", key.c_str(), path.c_str(),
                         out.data.size(), out.rows, out.cols);
    return PF_OK;
}

int yml_read_scalar(const std::string &path, const std::string &key, double &out)
{
    std::vector<std::string> lines;
    int st = read_lines(path, lines);
    if (st) return st;
    const int k = find_key(lines, key);
    if (k < 0) return set_error(PF_ERR_IO, "key '%s' not found in %s", key.c_str(), path.c_str());
    const std::string &l = lines[k];
    const std::string val = trim(l.substr(l.find(':') + 1));
    char *end = nullptr;
    out = std::strtod(val.c_str(), &end);
    if (end == val.c_str()) return set_error(PF_ERR_IO, "'%s' in %s is not a number", key.c_str(), path.c_str());
    return PF_OK;
}

}  // namespace mkfbt

using namespace mkfbt;

extern "C" {

int pf_yml_read_matrix(const char *path, const char *key, int32_t *rows, int32_t *cols, double *out,
                       int64_t cap)
{
    if (!path || !key) return set_error(PF_ERR_ARG, "null path/key");
    YmlMatrix m;
    int st = yml_read_matrix(path, key, m);
    if (st) return st;
    if (rows) *rows = m.rows;
    if (cols) *cols = m.cols;
    if (out) {
        if ((int64_t)m.data.size() > cap) return set_error(PF_ERR_ARG, "output capacity %lld too small", (long long)cap);
        for (size_t i = 0; i < m.data.size(); ++i) out[i] = m.data[i];
    }
    return PF_OK;
}

int pf_read_camera_info(const char *path, int32_t *width, int32_t *height, double *K, double *D, double *P)
{
    if (!path) return set_error(PF_ERR_ARG, "null path");
    double w = 0, hgt = 0;
    int st = yml_read_scalar(path, "image_width", w);
    if (st) return st;
    st = yml_read_scalar(path, "image_height", hgt);
    if (st) return st;
    YmlMatrix k, d, p;
    if ((st = yml_read_matrix(path, "camera_matrix", k))) return st;
    if ((st = yml_read_matrix(path, "distortion_coefficients", d))) return st;
    if ((st = yml_read_matrix(path, "projection_matrix", p))) return st;
    if (k.data.size() != 9 || p.data.size() != 12 || d.data.size() > 5)
        return set_error(PF_ERR_IO, "unexpected camera_info shapes in %s", path);
    if (width) *width = (int32_t)w;
    if (height) *height = (int32_t)hgt;
    if (K) for (int i = 0; i < 9; ++i) K[i] = k.data[i];
    if (D) for (int i = 0; i < 5; ++i) D[i] = i < (int)d.data.size() ? d.data[i] : 0.0;
    if (P) for (int i = 0; i < 12; ++i) P[i] = p.data[i];
    return PF_OK;
}

int pf_load_model_yml(pf_handle *h, const char *path, const char *gamma_path, int32_t m,
                      const int32_t *sel, const double *R)
{
    if (!h || !path) return set_error(PF_ERR_ARG, "null handle/path");
    YmlMatrix means, covs, weights, proj, mean, gamma;
    int st;
    if ((st = yml_read_matrix(path, "means", means))) return st;
    if ((st = yml_read_matrix(path, "covs", covs))) return st;
    if ((st = yml_read_matrix(path, "weights", weights))) return st;
    if ((st = yml_read_matrix(path, "pca_proj", proj))) return st;
    if ((st = yml_read_matrix(path, "pca_mean", mean))) return st;
    if ((st = yml_read_matrix(gamma_path ? gamma_path : path, "gamma", gamma))) return st;
    const int K = means.rows, d = means.cols;
    if (covs.rows != K * d || covs.cols != d) return set_error(PF_ERR_IO, "covs is %d x %d, expected %d x %d", covs.rows, covs.cols, K * d, d);
    if ((int)weights.data.size() != K) return set_error(PF_ERR_IO, "weights has %zu entries, expected %d", weights.data.size(), K);
    if ((int)gamma.data.size() < K) return set_error(PF_ERR_IO, "gamma has %zu entries, expected %d", gamma.data.size(), K);
    if (proj.rows != d) return set_error(PF_ERR_IO, "pca_proj has %d rows, expected %d", proj.rows, d);
    if ((int)mean.data.size() != proj.cols) return set_error(PF_ERR_IO, "pca_mean/pca_proj size mismatch");
    return pf_load_model(h, K, d, proj.cols, m, sel, means.data.data(), covs.data.data(),
                         weights.data.data(), gamma.data.data(), proj.data.data(), mean.data.data(), R);
}

}  // extern "C"
