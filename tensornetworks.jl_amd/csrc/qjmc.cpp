// Quantum-jump Monte Carlo trajectories run natively: the step loop of qjmc_simulation,
// src/algorithms/mps/qjmc.jl:28-167 (classical branch, single-site jump operators), with
// applygates! (src/structures/mps/gatelist.jl:191-223), the emission rates of
// qjmc_emission_rates (:170-220) and the jump of :103-107, over the device operators of
// this library. Trajectories are independent: worker threads each own a context (one HIP
// stream), so trajectories overlap on the GPU while every one of them consumes its own
// uniforms in the reference's order (one discarded draw, the jump test, the channel).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <mutex>
#include <thread>
#include <vector>

#include "tnh_internal.hpp"

using namespace tnh;

namespace {

void ck(int rc, tnh_ctx *ctx) {
  if (rc != TNH_OK) throw Error(rc, ctx->err);
}

struct Worker {
  tnh_ctx *ctx = nullptr;
  const tnh_qjmc_problem *p = nullptr;
  std::vector<tnh_buf *> gates;  // one device tensor per gate entry
  std::vector<tnh_buf *> sites;
  int center = 0;
  std::vector<double> esc_ops, obs_ops;  // d x d interleaved matrices, term after term
  std::vector<int> esc_first, obs_first, ones;

  ~Worker() {
    if (!ctx) return;
    for (tnh_buf *b : gates) tnh_buf_free(ctx, b);
    for (tnh_buf *b : sites) tnh_buf_free(ctx, b);
    tnh_ctx_destroy(ctx);
  }

  tnh_buf *upload_matrix(const double *interleaved, int ndim, const int64_t *dims, int64_t n) {
    bool cplx = false;
    for (int64_t i = 0; i < n; ++i) cplx |= interleaved[2 * i + 1] != 0.0;
    tnh_buf *b = nullptr;
    ck(tnh_buf_upload(ctx, interleaved, TNH_C128, cplx ? TNH_C128 : TNH_F64, ndim, dims, &b), ctx);
    return b;
  }

  void setup() {
    const int d = p->d;
    int ng = 0;
    for (int l = 0; l < p->nlayers; ++l) ng += p->layer_count[l];
    for (int g = 0; g < ng; ++g) {
      const double *src = p->gate_data + 2 * p->gate_offset[g];
      if (p->gate_range[g] == 2) {
        int64_t dims[4] = {d, d, d, d};
        gates.push_back(upload_matrix(src, 4, dims, int64_t(d) * d * d * d));
      } else {
        int64_t dims[2] = {d, d};
        gates.push_back(upload_matrix(src, 2, dims, int64_t(d) * d));
      }
    }
    // escape operators c^2 L^H L of the jump terms (qjmc.jl:11-20), as d x d matrices
    for (int t = 0; t < p->njump; ++t) {
      const double *L = p->jump_op + 2 * size_t(d) * d * t;
      for (int j = 0; j < d; ++j)
        for (int i = 0; i < d; ++i) {  // (L^H L)[i, j] = sum_k conj(L[k, i]) L[k, j]
          std::complex<double> s = 0;
          for (int k = 0; k < d; ++k)
            s += std::conj(std::complex<double>(L[2 * (k + d * i)], L[2 * (k + d * i) + 1])) *
                 std::complex<double>(L[2 * (k + d * j)], L[2 * (k + d * j) + 1]);
          esc_ops.push_back(s.real());
          esc_ops.push_back(s.imag());
        }
      esc_first.push_back(p->jump_site[t]);
    }
    for (int o = 0; o < p->nobs; ++o) {
      obs_ops.insert(obs_ops.end(), p->obs_op + 2 * size_t(d) * d * o, p->obs_op + 2 * size_t(d) * d * (o + 1));
      obs_first.push_back(p->obs_site[o]);
    }
    ones.assign(std::max(p->njump, p->nobs) + 1, 1);
  }

  void reset_state() {
    for (tnh_buf *b : sites) tnh_buf_free(ctx, b);
    sites.clear();
    const int d = p->d;
    for (int i = 0; i < p->N; ++i) {
      int64_t dims[3] = {1, d, 1};
      sites.push_back(upload_matrix(p->init_state + 2 * size_t(d) * i, 3, dims, d));
    }
    center = 0;
  }

  void movecenter(int idx, double cutoff = 0.0, int maxdim = 0, int mindim = 1) {
    ck(tnh_mps_movecenter(ctx, p->N, sites.data(), center, idx, cutoff, maxdim, mindim), ctx);
    center = idx;
  }

  double norm() {
    if (center == 0) movecenter(1);
    double v = 0;
    ck(tnh_buf_norm(ctx, sites[center - 1], &v), ctx);
    return v;
  }

  void normalize() {
    const double n = norm();
    ck(tnh_buf_scale(ctx, sites[center - 1], 1.0 / n, 0.0), ctx);
  }

  // applygate(psi, site, gate, direction; ...), gatelist.jl:137-171, as the host mirror does
  void applygate(int site, int g, bool direction) {
    const int N = p->N;
    if (p->gate_range[g] == 2) {
      int k = 0;
      double disc = 0;
      ck(tnh_applygate2(ctx, &sites[site - 1], &sites[site], gates[g], direction ? 1 : 0, p->cutoff,
                        p->maxdim, p->mindim, &k, &disc),
         ctx);
      center = direction ? site : site + 1;
    } else {
      ck(tnh_apply_site_op(ctx, &sites[site - 1], gates[g]), ctx);
      center = site;
      const int nxt = site + 1 - 2 * (direction ? 1 : 0);
      if (nxt > 0 && nxt <= N) movecenter(nxt);
    }
  }

  // applygates!(psi, gates; cutoff, maxdim, mindim), gatelist.jl:191-223
  void applygates() {
    int base = 0;
    for (int l = 0; l < p->nlayers; ++l) {
      const int n = p->layer_count[l];
      if (n == 0) continue;
      const int first = p->gate_site[base];
      const int last = p->gate_site[base + n - 1] + p->gate_range[base + n - 1] - 1;
      const bool direction = !(std::abs(center - first) < std::abs(center - last));
      for (int i = 1; i <= n; ++i) {
        const int gi = direction ? n + 1 - i : i;
        const int g = base + gi - 1;
        const int ctr = direction ? p->gate_site[g] + p->gate_range[g] - 1 : p->gate_site[g];
        movecenter(ctr, p->cutoff, p->maxdim, p->mindim);
        applygate(p->gate_site[g], g, direction);
      }
      base += n;
    }
  }

  void inner(int nterms, const int *first, const double *ops, double *out) {
    ck(tnh_inner_oplist(ctx, p->N, sites.data(), nterms, first, ones.data(), ops, out), ctx);
  }

  int maxbond() const {
    int64_t m = 0;
    for (size_t i = 0; i + 1 < sites.size(); ++i) m = std::max(m, sites[i]->dims[2]);
    return int(m);
  }

  // one trajectory (qjmc.jl:53-164)
  void run(const double *U, int64_t nu, double *meas, int *njumps, int *jumps, int *jump_steps, int maxjumps,
           int *maxb) {
    reset_state();
    int64_t pos = 0;
    auto draw = [&]() {
      TNH_REQUIRE(pos < nu, TNH_ERR_ARG, "qjmc: not enough uniforms for a trajectory");
      return U[pos++];
    };
    int save = 0, nj = 0;
    std::vector<double> out(2 * std::max(1, std::max(p->njump, p->nobs)));
    auto measure = [&]() {
      if (p->nobs > 0) {
        inner(p->nobs, obs_first.data(), obs_ops.data(), out.data());
        std::copy(out.begin(), out.begin() + 2 * p->nobs, meas + size_t(save) * 2 * p->nobs);
      }
      ++save;
    };
    measure();
    std::vector<double> rates(p->njump);
    for (int step = 1; step <= p->nsteps; ++step) {
      applygates();
      (void)draw();  // qjmc.jl:64: drawn and unused on the classical branch
      normalize();
      inner(p->njump, esc_first.data(), esc_ops.data(), out.data());
      double er = 0;
      for (int t = 0; t < p->njump; ++t) {
        const double c = p->jump_coeff[t];
        rates[t] = std::abs(std::complex<double>(c * c * out[2 * t], c * c * out[2 * t + 1]));
        er += rates[t];
      }
      const double prob = std::exp(-er * p->dt);
      if (draw() > prob) {
        const double r = draw();
        int idx = -1;
        double cum = 0;
        for (int t = 0; t < p->njump && idx < 0; ++t) {
          cum += rates[t];
          if (r < cum / er) idx = t;
        }
        TNH_REQUIRE(idx >= 0, TNH_ERR_NUMERIC, "qjmc: no jump channel selected");
        movecenter(1);
        int64_t dims[2] = {p->d, p->d};
        tnh_buf *L = upload_matrix(p->jump_op + 2 * size_t(p->d) * p->d * idx, 2, dims, int64_t(p->d) * p->d);
        int rc = tnh_apply_site_op(ctx, &sites[p->jump_site[idx] - 1], L);
        tnh_buf_free(ctx, L);
        ck(rc, ctx);
        movecenter(p->N);
        movecenter(1, p->cutoff, p->maxdim, p->mindim);
        normalize();
        if (nj < maxjumps) {
          jumps[nj] = idx + 1;
          jump_steps[nj] = step;
        }
        ++nj;
      }
      if (step % p->savesteps == 0) measure();
    }
    *njumps = nj;
    *maxb = maxbond();
  }
};

}  // namespace

extern "C" int tnh_qjmc_run(tnh_ctx *ctx, const tnh_qjmc_problem *p, int ntraj, const double *uniforms,
                            int64_t nuniform, int nworkers, double *meas, int *njumps, int *jumps,
                            int *jump_steps, int maxjumps, int *maxbond) {
  if (!ctx) return TNH_ERR_ARG;
  return tnh_guard(ctx, [&] {
    TNH_REQUIRE(p && ntraj >= 0 && uniforms && njumps && maxbond && nworkers >= 1, TNH_ERR_ARG,
                "qjmc: bad arguments");
    TNH_REQUIRE(p->N >= 2 && p->d >= 1 && p->nsteps >= 0 && p->savesteps >= 1 && p->dt > 0, TNH_ERR_ARG,
                "qjmc: bad problem sizes");
    TNH_REQUIRE(p->nobs == 0 || meas, TNH_ERR_ARG, "qjmc: meas is NULL");
    TNH_REQUIRE(maxjumps == 0 || (jumps && jump_steps), TNH_ERR_ARG, "qjmc: jump arrays are NULL");
    for (int t = 0; t < p->njump; ++t)
      TNH_REQUIRE(p->jump_site[t] >= 1 && p->jump_site[t] <= p->N, TNH_ERR_ARG, "qjmc: jump site out of range");
    for (int o = 0; o < p->nobs; ++o)
      TNH_REQUIRE(p->obs_site[o] >= 1 && p->obs_site[o] <= p->N, TNH_ERR_ARG, "qjmc: observable site out of range");
    int ng = 0;
    for (int l = 0; l < p->nlayers; ++l) ng += p->layer_count[l];
    for (int g = 0; g < ng; ++g)
      TNH_REQUIRE((p->gate_range[g] == 1 || p->gate_range[g] == 2) && p->gate_site[g] >= 1 &&
                      p->gate_site[g] + p->gate_range[g] - 1 <= p->N,
                  TNH_ERR_ARG, "qjmc: gate out of range");
    const int nsave = p->nsteps / p->savesteps + 1;
    const int nw = std::min(nworkers, std::max(1, ntraj));
    std::atomic<int> next{0};
    std::atomic<bool> failed{false};
    std::mutex err_mu;
    int err_code = TNH_OK;
    std::string err_msg;
    auto work = [&]() {
      Worker w;
      try {
        const int rc = tnh_ctx_create(ctx->device, &w.ctx);
        if (rc != TNH_OK) throw Error(rc, "qjmc: creating a worker context failed");
        w.p = p;
        w.setup();
        for (int t = next++; t < ntraj && !failed; t = next++) {
          w.run(uniforms + size_t(t) * nuniform, nuniform,
                meas ? meas + size_t(t) * nsave * 2 * p->nobs : nullptr, njumps + t,
                jumps ? jumps + size_t(t) * maxjumps : nullptr,
                jump_steps ? jump_steps + size_t(t) * maxjumps : nullptr, maxjumps, maxbond + t);
        }
      } catch (const Error &e) {
        std::lock_guard<std::mutex> g(err_mu);
        if (!failed.exchange(true)) {
          err_code = e.code;
          err_msg = e.msg;
        }
      } catch (const std::exception &e) {
        std::lock_guard<std::mutex> g(err_mu);
        if (!failed.exchange(true)) {
          err_code = TNH_ERR_HIP;
          err_msg = e.what();
        }
      }
    };
    std::vector<std::thread> threads;
    for (int i = 0; i < nw; ++i) threads.emplace_back(work);
    for (auto &th : threads) th.join();
    TNH_REQUIRE(!failed, err_code, "qjmc worker: " + err_msg);
  });
}
