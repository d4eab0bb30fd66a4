// KF_model.hpp -- mirror of the reference KF_model (src/KF_model.h:8-16).
#pragma once

#include "Mat.hpp"

namespace mkfbt {
namespace pfrao {

// Linear Kalman filter model with public Q, R, F, B, H, BH.  predict/update
// run on the GPU (pf_kf_predict / pf_kf_update).
class KF_model {
public:
    KF_model();
    ~KF_model();
    Mat Q, R, F, B, H, BH;
    // state = F state + B ; cov = F cov F^T + Q  (KF_model.cpp:11-15)
    void predict(Mat &state, Mat &cov);
    // Kalman update with the measurement (KF_model.cpp:17-25)
    void update(Mat measurement, Mat &state, Mat &cov);
};

}  // namespace pfrao
}  // namespace mkfbt
