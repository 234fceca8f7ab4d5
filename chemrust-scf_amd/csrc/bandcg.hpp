// Batched band-CG refinement (bandcg.hip).
#pragma once
#include "impl.hpp"

namespace chs {

int bcg_scale_rows(chs_context* ctx, z64* w, int P, int n_b, const double* q);

struct BandCg {
  chs_context* ctx = nullptr;
  int n_pw = 0, n_b = 0;
  bool use_r = false;
  // lower-band S-orthogonalisation with the complex overlap (the reference,
  // cg_helpers.rs:188, keeps its real part only)
  bool full_overlap = false;
  DevBuf vecs;  // 12 band blocks
  DevBuf scal;  // per-band scalars, states, the n_b x n_b overlap
  DevBuf tpa;   // t(g)
  DevBuf rmat;  // R (P x P)
  int ensure_ws(chs_context* c, int n_pw, int n_b);
  int setup_precond(bool uspp);
  int precond(const z64* R, z64* G, const float* mlp, const void* st);
  int run(z64* psi, int max_steps, double tol, const float* mlp, double* eps_out, int* steps_out,
          int* conv_out);
  z64* vec(int i) { return (z64*)vecs.p + (size_t)i * n_pw * n_b; }
  ~BandCg();
};

}  // namespace chs
