// pf_model.hpp -- host-side model preparation shared by the single-filter
// and the batched C ABI: validation of the my_gmm inputs, the state basis in
// which the measurement matrix is a row selection, and the LDS model image
// of the step kernel.  (Reference: src/my_gmm.cpp:45-75, src/pfPose.cpp:61-65.)
#pragma once

#include <cstdint>
#include <vector>

namespace mkfbt {

struct HostModel {
    int D = 0, K = 0, m = 0, J = 0, npose = 0;
    std::vector<double> mu, sig, pi, gamma, H, BH, R;  // K x D, K x D x D, K, K, m x D, m, m x m
    double wmax = 0.0;
};

struct ModelImage {
    std::vector<float> blob;        // image: thresholds (4K floats), H, BH, R, then the component rows
    std::vector<float> mu, sg;      // reset table in the state basis: K x D, K x D(D+1)/2
    std::vector<double> T, Ti;      // state basis x' = T x and its inverse (D x D)
    bool sel_mode = false;          // H is a row selection in the state basis
    unsigned long long pi_total = 0;
    int head_floats = 0;            // thresholds + H, BH, R (always staged in LDS)
    bool comp_in_lds = true;        // the component rows fit the LDS budget with the head
};

// Validates the arguments of my_gmm::loadGaussian x K and fills hm (H = H1
// pca_proj^T, BH = H1 pca_mean^T, R defaulting to 100 I).  Returns a PF_*
// status (message set on error).
int host_model_from_args(HostModel &hm, int K, int d, int npose, int m, const int32_t *sel,
                         const double *means, const double *covs, const double *weights,
                         const double *gamma, const double *pca_proj, const double *pca_mean,
                         const double *R);

// Joint processing order (certain joints first, then cluttered ones, each in
// index order), 4 bits per slot.
unsigned joint_order(unsigned clutter_mask, int J);

// Builds the model image for N particles and the given processing order.
int make_model_image(const HostModel &hm, unsigned order_packed, int N, ModelImage &img);

}  // namespace mkfbt
