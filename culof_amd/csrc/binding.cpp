// pybind11 module `_culof`: the reference's Python-facing native surface
// (src/cuda/binding.cpp:23-56) over libculof_amd's C++ API.
//
// lof(points, k, normalize) coerces to a C-contiguous float32 array (so strided
// or Fortran-ordered views are read correctly), rejects anything but 2-D,
// releases the GIL for the GPU work, and returns a fresh 1-D float32 array.
// CudaError is registered as a RuntimeError subclass under the same name.

#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <algorithm>
#include <stdexcept>
#include <vector>

#include "culof.h"

namespace py = pybind11;

namespace {

using Array = py::array_t<float, py::array::c_style | py::array::forcecast>;

py::array_t<float> lof_py(const Array& points, int k, bool normalize) {
    const py::buffer_info buf = points.request();
    if (buf.ndim != 2) {
        throw std::invalid_argument("expected a 2-D array of shape (n_samples, n_features)");
    }
    std::vector<float> scores;
    {
        py::gil_scoped_release release;
        scores = culof::lof(static_cast<const float*>(buf.ptr), static_cast<int>(buf.shape[0]),
                            static_cast<int>(buf.shape[1]), k, normalize);
    }
    py::array_t<float> out(static_cast<py::ssize_t>(scores.size()));
    std::copy(scores.begin(), scores.end(), out.mutable_data());
    return out;
}

}  // namespace

PYBIND11_MODULE(_culof, m) {
    m.doc() = "culof native extension (HIP, gfx950)";
    m.def("lof", &lof_py, py::arg("points"), py::arg("k"), py::arg("normalize"),
          "Raw LOF scores for a (n_samples, n_features) array.");
    m.def("cuda_available", &culof::cuda_available);
    m.def("device_info", &culof::device_info);
    py::register_exception<culof::CudaError>(m, "CudaError", PyExc_RuntimeError);
}
