// Launchers of the contraction kernels (host side).
#pragma once

#include <hip/hip_runtime.h>

#include "context.hpp"

namespace hlbl {

void upload_terms_fp64(const qed::TermTable& t);
void upload_terms_fp32(const qed::TermTable& t);
void upload_terms_qed(const qed::TermTable& t);

// pass 1 of the flavours in flav_mask; leaves dz = [2][dzu 6x144 | dzsu 4x144] complex
void launch_pass1(hlbl_contractor* c, const double* S, const double* Y, hlbl_coord proc,
                  hlbl_coord gsx, hlbl_geom global, hlbl_geom local, int flav_mask,
                  hipStream_t st);
// gdz = gamma_nu gamma_5 dz, both flavours, reference layout per flavour
void launch_gamma(const double* dz, double* gdz, hipStream_t st);
// one flavour, from the caller's dzu [6][12][12] / dzsu [4][12][12] into g_dzu [6][4][12][12] /
// g_dzsu [4][4][12][12] (complex)
void launch_gamma_one(const double* dzu, const double* dzsu, double* g_dzu, double* g_dzsu, hipStream_t st);
// pass-2 B operands from gdz
void launch_bmat(const double* gdz, double* bmat, hipStream_t st);
// pass 2; out[f] (or out[0] with single_flavour) = or += the kernel sum
void launch_pass2(hlbl_contractor* c, const double* S, const double* Y, const double* bmat,
                  hlbl_coord proc, hlbl_coord gsx, hlbl_coord yv, double xunit, hlbl_geom global,
                  hlbl_geom local, int flav_mask, double* out, int accumulate, int single_flavour,
                  hipStream_t st);

// dst[i] += src[i]
void launch_add(double* dst, const double* src, size_t n, hipStream_t st);
// the 2+2 piece P1 of one block (flavour f) for source w into d_p1 [4][4][4][lmax]
void launch_p1_2p2(hlbl_contractor* c, const double* blk, int f, hlbl_coord gsw, double* d_p1,
                   hipStream_t st);
// the 2+2 pieces P2_0, P2_1, P3, P4_0, P4_1 from Pi [V][16] for each target (global
// sites, host array) into d_p234 [n_targets][5][4][4][4]
void launch_p234_2p2(hlbl_contractor* c, const double* pi, hlbl_coord gsw, const hlbl_coord* targets,
                     int n_targets, double xunit, double* d_p234, hipStream_t st);

// device random fill of one block
void launch_fill_random(double* d_out, hlbl_geom local, hlbl_coord proc, uint64_t seed,
                        hipStream_t st);

// QED kernel at n argument pairs (test entry) and channel interpolation; the point
// kernel with the caller's scratch of n * 3 * 384 doubles
void launch_qed_points(const double* table, const double* x, const double* y, int n, double* kerv,
                       double* scratch, hipStream_t st);
void launch_qed_kernel(const hlbl_qed* q, int mode, const double* x, const double* y, int n,
                       double* kerv, hipStream_t st);
void launch_qed_channels(const hlbl_qed* q, int mode, const double* acb, int n, double* out,
                         hipStream_t st);

}  // namespace hlbl
