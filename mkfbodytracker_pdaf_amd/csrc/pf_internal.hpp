// pf_internal.hpp -- host-side helpers shared by the C ABI translation units.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <string>
#include <vector>

namespace mkfbt {

// Records the message for pf_last_error() and returns code.
int set_error(int code, const char *fmt, ...);
// The C ABI status of a device status word (kStatus* bits of pf_common.hpp):
// PF_ERR_NOT_SPD (message set) if an innovation covariance was not positive
// definite, else the warning bits PF_WARN_DEGENERATE | PF_WARN_ILLCOND.
int status_code(unsigned word);

// Philox4x32-10 (Salmon et al., SC'11) stream of 32-bit words.
void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void philox_u32(uint64_t seed, uint32_t stream, uint64_t first, int64_t count, uint32_t *out);

// One matrix node of an OpenCV FileStorage / ROS camera_info YAML file.
struct YmlMatrix {
    int rows = 0, cols = 0;
    char dt = 'd';
    std::vector<double> data;
};
// Returns 0 and fills out, or a PF_ERR_* code (with the message set).
int yml_read_matrix(const std::string &path, const std::string &key, YmlMatrix &out);
int yml_read_scalar(const std::string &path, const std::string &key, double &out);

}  // namespace mkfbt
