// Device-resident two-site DMRG state and sweep: dmrg(psi, H::GMPS...),
// src/algorithms/mps/dmrg.jl:128-154 and the bond loop of :34-63, with the
// environment cache of ProjMPS / ProjMPSSum (src/structures/mps/projmps.jl:1-42,
// abstractprojmps.jl:60-81, projmpssum.jl:63-73). The convergence bookkeeping of
// dmrg.jl:66-95 stays with the host caller, which calls tnh_dmrg_sweep per pass.
#include <algorithm>
#include <cmath>
#include <vector>

#include "tnh_internal.hpp"

using namespace tnh;

struct tnh_dmrg {
  tnh_ctx *ctx = nullptr;
  int N = 0;
  int nterms = 0;
  std::vector<tnh_buf *> psi;                 // N site tensors (F64 or C128)
  std::vector<std::vector<tnh_buf *>> W;      // [term][site]
  std::vector<double> coeffs;
  std::vector<std::vector<tnh_buf *>> blocks; // [term][site] environment blocks
  std::vector<tnh_buf *> shadowL, shadowR;    // FP32 copies of the bond's blocks, per term
  tnh_buf *theta = nullptr;
  tnh_buf *tmp = nullptr;
  // Per bond and sweep direction, the eigenvectors of the last split's Gram matrix there:
  // the next split of that bond in the same direction starts its Jacobi from them
  // (p x p FP64 each, p = 2 chi: 134 MB per bond at chi = 2048).
  std::vector<tnh_buf *> warm[2];
  int env_center = 0;   // ProjMPS centre (0 = blocks not built)
  int psi_center = 0;
  bool cplx = false;    // psi (and every term) held complex
  double ritz_re = 0, ritz_im = 0;  // the last Ritz value of the last pass
  // Rank-1 terms coeff |phi><phi|, ProjMPS(phi, psi; rank=2, squared=true, coeff)
  // (dmrg.jl:144-145): phi's site tensors, the overlap blocks per site, and the projection of
  // phi onto the bond being updated.
  struct Proj {
    std::vector<tnh_buf *> phi, blocks;
    double coeff = 1.0;
    tnh_buf *P = nullptr;
  };
  std::vector<Proj> proj;
  tnh_buf *one = nullptr;  // the 1 x 1 overlap block beyond the chain ends (edgeblock(2))

  void free_proj(Proj &q) {
    for (auto &b : q.phi) buf_delete(ctx, b);
    for (auto &b : q.blocks) buf_delete(ctx, b);
    buf_delete(ctx, q.P);
    q.phi.clear();
    q.blocks.clear();
    q.P = nullptr;
  }

  ~tnh_dmrg() {
    auto del = [&](tnh_buf *&b) {
      buf_delete(ctx, b);
      b = nullptr;
    };
    for (auto &b : psi) del(b);
    for (auto &v : W)
      for (auto &b : v) del(b);
    for (auto &v : blocks)
      for (auto &b : v) del(b);
    for (auto &b : shadowL) del(b);
    for (auto &b : shadowR) del(b);
    del(theta);
    del(tmp);
    for (auto &v : warm)
      for (auto &b : v) del(b);
    for (auto &q : proj) free_proj(q);
    del(one);
  }

  // block(projV, idx), abstractprojmps.jl:33-46 (nullptr = edge block)
  const tnh_buf *block(int t, int idx) const {
    if (idx < 1 || idx > N) return nullptr;
    return blocks[t][idx - 1];
  }

  const tnh_buf *pblock(const Proj &q, int idx) const {
    if (idx < 1 || idx > N) return one;
    return q.blocks[idx - 1];
  }

  void buildleft(int idx, int prec) {
    for (int t = 0; t < nterms; ++t)
      env_left(ctx, block(t, idx - 1), psi[idx - 1], W[t][idx - 1], &blocks[t][idx - 1], prec);
    for (auto &q : proj)  // overlap blocks stay FP64 in every pass
      proj_env_left(ctx, pblock(q, idx - 1), q.phi[idx - 1], psi[idx - 1], &q.blocks[idx - 1]);
  }
  void buildright(int idx, int prec) {
    for (int t = 0; t < nterms; ++t)
      env_right(ctx, block(t, idx + 1), psi[idx - 1], W[t][idx - 1], &blocks[t][idx - 1], prec);
    for (auto &q : proj)
      proj_env_right(ctx, pblock(q, idx + 1), q.phi[idx - 1], psi[idx - 1], &q.blocks[idx - 1]);
  }

  // movecenter!(projV, idx), abstractprojmps.jl:60-81
  void env_movecenter(int idx, int prec) {
    if (env_center == 0) {
      for (int i = 1; i < idx; ++i) buildleft(i, prec);
      for (int i = 1; i <= N - idx; ++i) buildright(N + 1 - i, prec);
    } else if (idx > env_center) {
      for (int i = 1; i <= idx - env_center; ++i) buildleft(env_center - 1 + i, prec);
    } else if (idx < env_center) {
      for (int i = 1; i <= env_center - idx; ++i) buildright(env_center + 1 - i, prec);
    }
    env_center = idx;
  }
};

namespace {

// the local eigensolve of dmrg.jl:51-53: Lanczos, or Arnoldi for ishermitian = false
double local_eig(tnh_dmrg *st, const Matvec &mv, const tnh_dmrg_params *p, int *nmv) {
  if (p->nonhermitian) {
    const std::complex<double> l = eigsolve_nonherm_op(st->ctx, mv, st->theta, p->krylovdim,
                                                       p->maxiter, p->lanczos_tol, nmv);
    st->ritz_re = l.real();
    st->ritz_im = l.imag();
    return l.real();
  }
  st->ritz_re = eigsolve_ground_op(st->ctx, mv, st->theta, p->krylovdim, p->maxiter,
                                   p->lanczos_tol, nmv);
  st->ritz_im = 0.0;
  return st->ritz_re;
}

void copy_in(tnh_ctx *ctx, const tnh_buf *src, bool cplx, tnh_buf **dst) {
  TNH_REQUIRE(src, TNH_ERR_ARG, "NULL tensor");
  const int dt = cplx ? TNH_C128 : TNH_F64;
  buf_ensure(ctx, dst, dt, src->ndim, src->dims);
  if (src->numel == 0) return;
  if (src->dtype == dt) {
    copy_device(ctx, (*dst)->ptr, src->ptr, src->bytes);
  } else {
    convert(ctx, src->ptr, src->dtype, (*dst)->ptr, dt, src->numel);
  }
}

// y = sum_t coeff_t H_t x + sum_q coeff_q |phi_q><phi_q| x for the bond (site1, site1 + 1),
// product(::ProjMPSSum), projmpssum.jl:63-73 (the rank-1 terms after the MPO terms)
void bond_matvec(tnh_dmrg *st, int site1, int prec, const tnh_buf *x, tnh_buf *y) {
  tnh_ctx *ctx = st->ctx;
  if (st->nterms == 0) fill(ctx, y->ptr, y->dtype, y->numel, 0.0);
  for (int t = 0; t < st->nterms; ++t) {
    const tnh_buf *L = st->block(t, site1 - 1);
    const tnh_buf *R = st->block(t, site1 + 2);
    if (prec != TNH_PREC_F64) {
      if (L) L = st->shadowL[t];
      if (R) R = st->shadowR[t];
    }
    if (t == 0) {
      heff_apply(ctx, L, st->W[t][site1 - 1], st->W[t][site1], R, x, y, st->coeffs[t], 0.0, prec);
    } else {
      buf_ensure(ctx, &st->tmp, y->dtype, y->ndim, y->dims);
      heff_apply(ctx, L, st->W[t][site1 - 1], st->W[t][site1], R, x, st->tmp, st->coeffs[t], 0.0,
                 prec);
      axpby(ctx, y->numel * (is_complex(y->dtype) ? 2 : 1), 1.0,
            static_cast<const double *>(st->tmp->ptr), 1.0, static_cast<double *>(y->ptr));
    }
  }
  for (const auto &q : st->proj) proj_apply(ctx, q.P, x, y, q.coeff);
}

// y = sum_t coeff_t H_t x + the rank-1 terms for the single site `site` (nsites = 1)
void site_matvec(tnh_dmrg *st, int site, int prec, const tnh_buf *x, tnh_buf *y) {
  tnh_ctx *ctx = st->ctx;
  if (st->nterms == 0) fill(ctx, y->ptr, y->dtype, y->numel, 0.0);
  for (int t = 0; t < st->nterms; ++t) {
    const tnh_buf *L = st->block(t, site - 1);
    const tnh_buf *R = st->block(t, site + 1);
    if (prec != TNH_PREC_F64) {
      if (L) L = st->shadowL[t];
      if (R) R = st->shadowR[t];
    }
    if (t == 0) {
      heff1_apply(ctx, L, st->W[t][site - 1], R, x, y, st->coeffs[t], prec);
    } else {
      buf_ensure(ctx, &st->tmp, y->dtype, y->ndim, y->dims);
      heff1_apply(ctx, L, st->W[t][site - 1], R, x, st->tmp, st->coeffs[t], prec);
      axpby(ctx, y->numel * (is_complex(y->dtype) ? 2 : 1), 1.0,
            static_cast<const double *>(st->tmp->ptr), 1.0, static_cast<double *>(y->ptr));
    }
  }
  for (const auto &q : st->proj) proj_apply(ctx, q.P, x, y, q.coeff);
}

// FP32 copies of the blocks left of site `lo` and right of site `hi` (the sites updated)
void make_shadows(tnh_dmrg *st, int lo, int hi) {
  tnh_ctx *ctx = st->ctx;
  for (int t = 0; t < st->nterms; ++t) {
    const tnh_buf *L = st->block(t, lo - 1);
    const tnh_buf *R = st->block(t, hi + 1);
    if (L) {
      const int dt = is_complex(L->dtype) ? TNH_C64 : TNH_F32;
      buf_ensure(ctx, &st->shadowL[t], dt, L->ndim, L->dims);
      convert(ctx, L->ptr, L->dtype, st->shadowL[t]->ptr, dt, L->numel);
    }
    if (R) {
      const int dt = is_complex(R->dtype) ? TNH_C64 : TNH_F32;
      buf_ensure(ctx, &st->shadowR[t], dt, R->ndim, R->dims);
      convert(ctx, R->ptr, R->dtype, st->shadowR[t]->ptr, dt, R->numel);
    }
  }
}

// One pass of the one-site update (dmrg.jl:35-63 with nsites = 1): at each site the
// environments move there, the site tensor is the Krylov start, the Ritz vector replaces
// it, and replacesites!'s one-site branch (gmps.jl:204-213) moves the centre one site on
// with an untruncated split and normalises the new centre (at the last site of the pass
// the centre stays and that tensor is normalised). Returns the last Ritz value.
double one_site_pass(tnh_dmrg *st, int direction, const tnh_dmrg_params *p) {
  tnh_ctx *ctx = st->ctx;
  const int N = st->N, prec = p->precision;
  double cost = 0;
  for (int j = 1; j <= N; ++j) {
    const int site = direction ? N + 1 - j : j;
    st->env_movecenter(site, prec);  // :41
    tnh_buf *A = st->psi[site - 1];
    const int64_t td[4] = {A->dims[0], A->dims[1], 1, A->dims[2]};  // the Krylov solver's view
    buf_ensure(ctx, &st->theta, A->dtype, 4, td);
    copy_device(ctx, st->theta->ptr, A->ptr, A->bytes);
    if (prec != TNH_PREC_F64) make_shadows(st, site, site);
    for (auto &q : st->proj)
      proj_project1(ctx, st->pblock(q, site - 1), q.phi[site - 1], st->pblock(q, site + 1), &q.P);
    int nmv = 0;
    cost = local_eig(
        st, [&](const tnh_buf *x, tnh_buf *y) { site_matvec(st, site, prec, x, y); }, p, &nmv);
    copy_device(ctx, A->ptr, st->theta->ptr, A->bytes);  // psi[site] = A, gmps.jl:205
    const int nxt = site + 1 - 2 * direction;
    int c = site;
    if (nxt >= 1 && nxt <= N) {  // movecenter!(psi, site + 1 - 2 direction), :206-208
      if (direction)
        mps_moveleft(ctx, &st->psi[nxt - 1], &st->psi[site - 1], 0.0, 0, 1);
      else
        mps_moveright(ctx, &st->psi[site - 1], &st->psi[nxt - 1], 0.0, 0, 1);
      c = nxt;
    }
    st->psi_center = c;
    // normalize!(psi), gmps.jl:46-51: the centre tensor divided by its norm
    tnh_buf *C = st->psi[c - 1];
    tnh_buf v = *C;
    v.dtype = TNH_F64;
    v.numel = C->numel * (is_complex(C->dtype) ? 2 : 1);
    double nr = 0, ni = 0;
    cdot_host(ctx, &v, &v, &nr, &ni);
    nr = std::sqrt(std::max(nr, 0.0));
    TNH_REQUIRE(std::isfinite(nr) && nr > 0, TNH_ERR_NUMERIC, "dmrg: zero centre tensor");
    scale(ctx, C->ptr, C->dtype, C->numel, 1.0 / nr);
  }
  st->env_movecenter(direction ? 1 : N, prec);  // :62
  return cost;
}

}  // namespace

extern "C" {

int tnh_dmrg_create(tnh_ctx *ctx, int N, const tnh_buf *const *psi, int nterms,
                    const tnh_buf *const *mpo, const double *coeffs, tnh_dmrg **out) {
  if (!ctx || !out) return TNH_ERR_ARG;
  *out = nullptr;
  tnh_dmrg *st = new (std::nothrow) tnh_dmrg();
  if (!st) return TNH_ERR_NOMEM;
  st->ctx = ctx;
  int rc = tnh_guard(ctx, [&] {
    TNH_REQUIRE(N >= 2, TNH_ERR_ARG, "two-site DMRG needs N >= 2");
    TNH_REQUIRE(nterms >= 0, TNH_ERR_ARG, "negative number of MPO terms");
    TNH_REQUIRE(psi && (mpo || nterms == 0), TNH_ERR_ARG, "NULL psi or mpo array");
    st->N = N;
    st->nterms = nterms;
    bool cplx = false;
    for (int i = 0; i < N; ++i) {
      TNH_REQUIRE(psi[i] && psi[i]->ndim == 3, TNH_ERR_SHAPE, "Psi must be a GMPS of rank 1 (vector).");
      cplx |= is_complex(psi[i]->dtype);
    }
    const int64_t d = psi[0]->dims[1];
    for (int t = 0; t < nterms; ++t)
      for (int i = 0; i < N; ++i) {
        const tnh_buf *w = mpo[size_t(t) * N + i];
        TNH_REQUIRE(w && w->ndim == 4, TNH_ERR_SHAPE,
                    "Hamiltonian must be composed of MPOs (rank 2) or MPSs (rank 1).");
        TNH_REQUIRE(w->dims[1] == d && w->dims[2] == d && psi[i]->dims[1] == d, TNH_ERR_SHAPE,
                    "GMPS must share the same properties.");
        cplx |= is_complex(w->dtype);
      }
    // bond consistency of psi and the MPOs
    for (int i = 0; i + 1 < N; ++i)
      TNH_REQUIRE(psi[i]->dims[2] == psi[i + 1]->dims[0], TNH_ERR_SHAPE, "psi bond dims do not match");
    TNH_REQUIRE(psi[0]->dims[0] == 1 && psi[N - 1]->dims[2] == 1, TNH_ERR_SHAPE,
                "psi edge bonds must be 1");
    for (int t = 0; t < nterms; ++t) {
      for (int i = 0; i + 1 < N; ++i)
        TNH_REQUIRE(mpo[size_t(t) * N + i]->dims[3] == mpo[size_t(t) * N + i + 1]->dims[0],
                    TNH_ERR_SHAPE, "MPO bond dims do not match");
      TNH_REQUIRE(mpo[size_t(t) * N]->dims[0] == 1 && mpo[size_t(t) * N + N - 1]->dims[3] == 1,
                  TNH_ERR_SHAPE, "MPO edge bonds must be 1");
    }
    st->cplx = cplx;
    st->psi.assign(N, nullptr);
    for (int i = 0; i < N; ++i) copy_in(ctx, psi[i], cplx, &st->psi[i]);
    st->W.assign(nterms, std::vector<tnh_buf *>(N, nullptr));
    st->blocks.assign(nterms, std::vector<tnh_buf *>(N, nullptr));
    st->shadowL.assign(nterms, nullptr);
    st->shadowR.assign(nterms, nullptr);
    st->coeffs.assign(nterms, 1.0);
    st->warm[0].assign(N, nullptr);
    st->warm[1].assign(N, nullptr);
    for (int t = 0; t < nterms; ++t) {
      if (coeffs) st->coeffs[t] = coeffs[t];
      for (int i = 0; i < N; ++i) copy_in(ctx, mpo[size_t(t) * N + i], cplx, &st->W[t][i]);
    }
    // movecenter!(psi, 1), dmrg.jl:133 (the caller's centre is unknown: canonicalise fully)
    mps_movecenter(ctx, N, st->psi.data(), 0, 1, 0.0, 0, 1);
    st->psi_center = 1;
    // ProjMPS(psi, H, psi; rank=2): centre 1 builds all right blocks, projmps.jl:38-40
    st->env_movecenter(1, TNH_PREC_F64);
    TNH_HIP(hipStreamSynchronize(ctx->stream));
  });
  if (rc != TNH_OK) {
    delete st;
    return rc;
  }
  *out = st;
  return TNH_OK;
}

int tnh_dmrg_add_projector(tnh_dmrg *st, const tnh_buf *const *phi, double coeff) {
  if (!st || !phi) return TNH_ERR_ARG;
  tnh_ctx *ctx = st->ctx;
  return tnh_guard(ctx, [&] {
    const int N = st->N;
    const int64_t d = st->psi[0]->dims[1];
    bool pc = false;
    for (int i = 0; i < N; ++i) {
      TNH_REQUIRE(phi[i] && phi[i]->ndim == 3, TNH_ERR_SHAPE,
                  "Hamiltonian must be composed of MPOs (rank 2) or MPSs (rank 1).");
      TNH_REQUIRE(phi[i]->dims[1] == d, TNH_ERR_SHAPE, "GMPS must share the same properties.");
      pc |= is_complex(phi[i]->dtype);
    }
    for (int i = 0; i + 1 < N; ++i)
      TNH_REQUIRE(phi[i]->dims[2] == phi[i + 1]->dims[0], TNH_ERR_SHAPE, "phi bond dims do not match");
    TNH_REQUIRE(phi[0]->dims[0] == 1 && phi[N - 1]->dims[2] == 1, TNH_ERR_SHAPE,
                "phi edge bonds must be 1");
    TNH_REQUIRE(!pc || st->cplx, TNH_ERR_ARG,
                "a complex rank-1 term needs a complex psi (create the state from complex tensors)");
    if (!st->one) {
      const int64_t od[2] = {1, 1};
      st->one = buf_new(ctx, TNH_F64, 2, od);
      fill(ctx, st->one->ptr, TNH_F64, 1, 1.0);
    }
    st->proj.emplace_back();
    tnh_dmrg::Proj &q = st->proj.back();
    q.coeff = coeff;
    q.phi.assign(N, nullptr);
    q.blocks.assign(N, nullptr);
    try {
      for (int i = 0; i < N; ++i) copy_in(ctx, phi[i], st->cplx, &q.phi[i]);
      // ProjMPS(...; center) builds every block on its side of the centre, projmps.jl:38-40
      const int c = std::max(st->env_center, 1);
      for (int i = 1; i < c; ++i)
        proj_env_left(ctx, st->pblock(q, i - 1), q.phi[i - 1], st->psi[i - 1], &q.blocks[i - 1]);
      for (int i = N; i > c; --i)
        proj_env_right(ctx, st->pblock(q, i + 1), q.phi[i - 1], st->psi[i - 1], &q.blocks[i - 1]);
      TNH_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
      st->free_proj(q);
      st->proj.pop_back();
      throw;
    }
  });
}

int tnh_dmrg_destroy(tnh_dmrg *st) {
  delete st;
  return TNH_OK;
}

// Stopping floor of the block-Jacobi Gram eigensolves in FP32 passes, relative to max_i G_ii
// (FP64 passes: 4 eps sqrt(p), ~6e-14 at p = 4096). An FP32 pass's theta is off by ~1e-6
// relative (its matvecs' rounding), so its Gram matrix is off by ~1e-6 max G_ii: resolving
// the eigenvectors below 1e-9 of it changes nothing that the FP64 passes after it keep
// (measured on the N = 100 TFIM at bond dimension 2048: the energy after one FP64 pass moved
// by 6e-13 per site, while the warm splits took 3.3 Jacobi sweeps instead of 6.3). Split-BF16x3
// passes keep the FP64 floor: their theta is noisier still, and with this floor their one-pass
// refinement came out worse in the same measurement.
constexpr double kF32GramFloor = 1e-9;

int tnh_dmrg_sweep(tnh_dmrg *st, int direction, const tnh_dmrg_params *p, double *energy,
                   int *maxbonddim, double *max_discarded) {
  if (!st || !p) return TNH_ERR_ARG;
  tnh_ctx *ctx = st->ctx;
  return tnh_guard(ctx, [&] {
    const int N = st->N;
    const int prec = p->precision;
    TNH_REQUIRE(prec >= TNH_PREC_F64 && prec <= TNH_PREC_BF16X3, TNH_ERR_ARG, "bad precision");
    TNH_REQUIRE(p->maxdim >= 0 && p->mindim >= 0, TNH_ERR_ARG, "maxdim, mindim must be >= 0");
    TNH_REQUIRE(st->nterms + st->proj.size() >= 1, TNH_ERR_ARG,
                "You must provide atleast one MPS/MPO for the Hamiltonian.");
    TNH_REQUIRE(p->nsites >= 0 && p->nsites <= 2, TNH_ERR_ARG, "nsites must be 1 or 2");
    TNH_REQUIRE(!p->nonhermitian || st->cplx, TNH_ERR_ARG,
                "ishermitian = false needs a complex state (create it from complex tensors)");
    double cost = 0, maxdisc = 0;
    if (p->nsites == 1) {
      cost = one_site_pass(st, direction, p);
      TNH_HIP(hipStreamSynchronize(ctx->stream));
      if (energy) *energy = cost;
      if (maxbonddim) {
        int D = 0;
        for (int i = 0; i + 1 < N; ++i) D = std::max<int>(D, int(st->psi[i]->dims[2]));
        *maxbonddim = D;
      }
      if (max_discarded) *max_discarded = 0.0;
      return;
    }
    for (int j = 1; j <= N - 1; ++j) {  // dmrg.jl:35
      const int site = direction ? N + 1 - j : j;
      const int site1 = direction ? site - 1 : site;
      // with profiling on, HIP events bracket the phases of the bond step (the interval
      // includes the host round trips inside a phase)
      hipEvent_t ph = ctx->prof_begin();
      st->env_movecenter(site, prec);  // :41
      ctx->prof_end("dmrg_env", ph, 0, 0);
      ph = ctx->prof_begin();
      // the psi centre must sit on the bond: it does after the previous split
      two_site_theta(ctx, st->psi[site1 - 1], st->psi[site1], &st->theta);  // :44-47
      if (prec != TNH_PREC_F64) make_shadows(st, site1, site1 + 1);
      for (auto &q : st->proj)  // project(projV, A, direction, 2), projmps.jl:152-179
        proj_project(ctx, st->pblock(q, site1 - 1), q.phi[site1 - 1], q.phi[site1],
                     st->pblock(q, site1 + 2), &q.P);
      int nmv = 0;
      cost = local_eig(
          st, [&](const tnh_buf *x, tnh_buf *y) { bond_matvec(st, site1, prec, x, y); }, p,
          &nmv);  // :50-55
      ctx->prof_end("dmrg_lanczos", ph, 0, 0);
      ph = ctx->prof_begin();
      int k = 0;
      double disc = 0;
      {
        // an FP32 pass stops its Gram eigensolves at a floor far below the error its theta
        // already carries (the FP64 passes that follow restore full accuracy)
        struct Floor {
          tnh_ctx *c;
          ~Floor() { c->gram_floor = 0.0; }
        } fl{ctx};
        ctx->gram_floor = prec == TNH_PREC_F32 ? kF32GramFloor : 0.0;
        replacesites2(ctx, st->theta, direction, 1, p->cutoff, p->maxdim, p->mindim,
                      &st->psi[site1 - 1], &st->psi[site1], &k, &disc, nullptr, 0, false,
                      &st->warm[direction ? 1 : 0][site1 - 1]);  // :57
      }
      ctx->prof_end("dmrg_split", ph, 0, 0);
      st->psi_center = direction ? site1 : site1 + 1;
      maxdisc = std::max(maxdisc, disc);
    }
    st->env_movecenter(direction ? 1 : N, prec);  // :62
    int D = 0;
    for (int i = 0; i + 1 < N; ++i) D = std::max<int>(D, int(st->psi[i]->dims[2]));
    TNH_HIP(hipStreamSynchronize(ctx->stream));
    if (energy) *energy = cost;
    if (maxbonddim) *maxbonddim = D;
    if (max_discarded) *max_discarded = maxdisc;
  });
}

int tnh_dmrg_energy(tnh_dmrg *st, int precision, double *energy) {
  if (!st) return TNH_ERR_ARG;
  (void)precision;
  return tnh_guard(st->ctx, [&] {
    // calculate(Hs) at the projection centre, projmps.jl:192-216 (summed over terms)
    double e = 0;
    const int c = st->env_center;
    TNH_REQUIRE(c >= 1 && c <= st->N, TNH_ERR_ARG, "environment centre not set");
    TNH_REQUIRE(st->psi_center == c, TNH_ERR_ARG,
                "energy: the MPS centre is not at the environment centre");
    for (int t = 0; t < st->nterms; ++t) {
      double im = 0;
      e += st->coeffs[t] * env_calculate(st->ctx, st->block(t, c - 1), st->psi[c - 1],
                                         st->W[t][c - 1], st->block(t, c + 1), &im);
    }
    // a rank-1 term's calculate is coeff <phi|psi> (projmps.jl:192-216 with objects [phi, psi])
    for (const auto &q : st->proj) {
      double re = 0, im = 0;
      proj_calculate(st->ctx, st->pblock(q, c - 1), q.phi[c - 1], st->psi[c - 1],
                     st->pblock(q, c + 1), &re, &im);
      e += q.coeff * re;
    }
    if (energy) *energy = e;
  });
}

int tnh_dmrg_last_ritz(tnh_dmrg *st, double *re, double *im) {
  if (!st) return TNH_ERR_ARG;
  if (re) *re = st->ritz_re;
  if (im) *im = st->ritz_im;
  return TNH_OK;
}

int tnh_dmrg_site(tnh_dmrg *st, int site, const tnh_buf **out) {
  if (!st || !out) return TNH_ERR_ARG;
  if (site < 1 || site > st->N) {
    st->ctx->err = "site out of range";
    return TNH_ERR_ARG;
  }
  *out = st->psi[site - 1];
  return TNH_OK;
}

int tnh_dmrg_center(tnh_dmrg *st, int *center) {
  if (!st || !center) return TNH_ERR_ARG;
  *center = st->psi_center;
  return TNH_OK;
}

int tnh_dmrg_bonddims(tnh_dmrg *st, int64_t *dims) {
  if (!st || !dims) return TNH_ERR_ARG;
  dims[0] = st->psi[0]->dims[0];
  for (int i = 0; i < st->N; ++i) dims[i + 1] = st->psi[i]->dims[2];
  return TNH_OK;
}

}  // extern "C"
