/**
 * culof - Local Outlier Factor on AMD Instinct MI355X (gfx950), C++ API.
 *
 * Source-compatible with the reference's public header
 * (include/culof.h:19-49): same namespace, same names, same signatures, same
 * exception types. `CudaError` keeps its reference name so existing callers
 * compile unchanged; here it reports HIP runtime failures.
 *
 * Scores match scikit-learn's `-LocalOutlierFactor(...).negative_outlier_factor_`:
 * about 1.0 for a point as densely surrounded as its neighbours, larger for one
 * in a sparser region.
 *
 * The C ABI underneath (culof_c.h) carries the same entry point for FFI
 * callers, plus the device-resident, precision-form and backward entry points.
 */

#ifndef CULOF_H
#define CULOF_H

#include <stdexcept>
#include <string>
#include <vector>

namespace culof {

/** Any HIP runtime failure (reference: include/culof.h:21-25). */
class CudaError : public std::runtime_error {
public:
    explicit CudaError(const std::string& what) : std::runtime_error(what) {}
};

/** True if a usable GPU is present (reference: include/culof.h:27-28). */
bool cuda_available();

/** Name, architecture, CU count and memory of the active device
 *  (reference: include/culof.h:30-31). */
std::string device_info();

/**
 * Compute Local Outlier Factor scores (reference: include/culof.h:33-47).
 *
 * @param points    host pointer to row-major (n_points x n_dims) float32 data
 * @param n_points  number of samples, >= 2
 * @param n_dims    number of features, >= 1
 * @param k         neighbours per point, 1 <= k <= n_points - 1
 * @param normalize z-score each feature before computing distances
 *
 * @throws std::invalid_argument for bad arguments, CudaError for device failure
 */
std::vector<float> lof(const float* points, int n_points, int n_dims, int k = 20,
                       bool normalize = false);

}  // namespace culof

#endif  // CULOF_H
