// Internal interface between the C-ABI layer (dgx_api.cpp) and the kernels (dgx_stage.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dgx_kernels.hpp"

namespace dgx {

bool supported_order(int d, int p);
int64_t stage_grid(int d, int p, int64_t n_elem);
hipError_t upload_ops(int p, const Ops1D& ops);
hipError_t launch_stage(int d, int p, int flux, const StageArgs& args, hipStream_t st);
hipError_t launch_trace(int d, int p, const double* u, double* tr, int64_t E, hipStream_t st);
hipError_t launch_ref_to_soa(const double* ref, double* soa, int64_t E, int nv, int ns, hipStream_t st);
hipError_t launch_soa_to_ref(const double* soa, double* ref, int64_t E, int nv, int ns, hipStream_t st);
hipError_t launch_reduce_max(const double* bm, int64_t nb, double* out, unsigned int* nonfinite, hipStream_t st);
hipError_t launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t st);

}  // namespace dgx
