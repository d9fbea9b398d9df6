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
// caller-layout src -> slot-order dst (times imass if given) and its face traces, in one pass
hipError_t launch_upload_trace(int d, int p, const double* src, int layout, const int32_t* slot_elem,
                               const double* imass, double* dst, double* tr, int64_t E, int64_t Ep, hipStream_t st);
hipError_t launch_to_internal(const double* src, int layout, double* dst, const int32_t* slot_elem, int64_t E,
                              int64_t Ep, int nv, int ns, hipStream_t st);
hipError_t launch_from_internal(const double* src, double* dst, int layout, const int32_t* elem_slot, int64_t E,
                                int64_t Ep, int nv, int ns, hipStream_t st);
hipError_t launch_hex3_stage(int flux, bool mixed, const Hex3Args& args, hipStream_t st);
hipError_t launch_hex3_trace(const double* u, double* tr, const int32_t* slot_elem, int64_t n_slots, hipStream_t st);
hipError_t launch_reduce_max(const double* bm, int64_t nb, double* out, unsigned int* nonfinite, hipStream_t st);
hipError_t launch_lin(int d, int p, int flux, bool adjoint, const LinArgs& args, hipStream_t st);
hipError_t launch_lin_block(int d, int p, int flux, const LinArgs& args, double* bt, hipStream_t st);
hipError_t launch_lin_block_f32(int d, int p, int flux, const LinArgs& args, float* bt, hipStream_t st);
hipError_t launch_scale_by_imass(const double* x, const double* imass, double* y, int64_t E, int nv, int ns,
                                 hipStream_t st);
hipError_t launch_ns_stage(int d, int p, int flux, const StageArgs& args, double* gtr, hipStream_t st);
int64_t ns_grid(int d, int p, int64_t n_elem);
// HEX p=4 Navier-Stokes: |J_f| n at every face point of every element, from the metric terms (set-up).
hipError_t launch_ns4_face_normals(const double* met, double* fnv, int64_t E, hipStream_t st);
// HEX p=4 Navier-Stokes: viscous normal-flux traces n.F_v of the state u (after an upload; stages write
// their own), [E][6][5][25].
hipError_t launch_ns4_vtrace(const double* u, const double* met, const double* imass, const double* fnv, double mu,
                             double Pr, double* vtr, int64_t E, hipStream_t st);
hipError_t launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t st);

}  // namespace dgx
