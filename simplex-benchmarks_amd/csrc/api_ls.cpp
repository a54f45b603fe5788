// C ABI of the learned smoother and the flexible CG that uses it
// (declared in include/simplexmf.h).
#include <cmath>
#include <cstring>

#include "api_handles.h"
#include "mlp.h"

using namespace smf;

struct mf_learned_smoother {
  std::unique_ptr<LearnedSmoother> ls;
};

namespace {
LearnedSmoother& get(mf_ls_t h) {
  if (!h || !h->ls) throw std::runtime_error("null learned smoother");
  return *h->ls;
}
}  // namespace

extern "C" {

void mf_ls_default_train_options(mf_train_options* opt) {
  TrainOptions d;
  opt->n_rhs = d.n_rhs;
  opt->pairs_per_rhs = d.pairs_per_rhs;
  opt->steps_per_pair = d.steps_per_pair;
  opt->lr = d.lr;
  opt->n_features = d.n_features;
  opt->corr_length = d.corr_length;
  opt->seed = d.seed;
  opt->loss = d.loss;
  opt->trajectory = d.trajectory;
}

void mf_ls_default_fcg_options(mf_fcg_options* opt) {
  FcgOptions d;
  opt->stall_ratio = d.stall_ratio;
  opt->force_fallback_at = d.force_fallback_at;
  opt->use_learned = d.use_learned;
}

int mf_ls_create(mf_op_t fine, uint64_t seed, mf_ls_t* out) {
  MF_TRY({
    if (!fine || !out) throw std::runtime_error("null argument");
    *out = nullptr;
    require_device(fine->op);
    auto h = std::make_unique<mf_learned_smoother>();
    h->ls = make_learned_smoother(fine->op, seed);
    *out = h.release();
  });
}

int mf_ls_destroy(mf_ls_t ls) {
  MF_TRY({ delete ls; });
}

int mf_ls_set_params(mf_ls_t h, const float* params, int64_t n) {
  MF_TRY({
    if (!params) throw std::runtime_error("null argument");
    if (n != kMlpParams) throw std::runtime_error("expected MF_LS_N_PARAMS parameters");
    ls_set_params(get(h), std::vector<float>(params, params + n));
  });
}

int mf_ls_get_params(mf_ls_t h, float* params, int64_t n) {
  MF_TRY({
    if (!params) throw std::runtime_error("null argument");
    if (n != kMlpParams) throw std::runtime_error("expected MF_LS_N_PARAMS parameters");
    const std::vector<float> p = ls_get_params(get(h));
    std::memcpy(params, p.data(), p.size() * sizeof(float));
  });
}

int mf_ls_save(mf_ls_t h, void* buf, uint64_t capacity, uint64_t* size) {
  MF_TRY({
    if (!size) throw std::runtime_error("null argument");
    const std::vector<uint8_t> blob = ls_serialize(get(h));
    *size = blob.size();
    if (buf) {
      if (capacity < blob.size()) throw std::runtime_error("buffer too small for the learned-smoother state");
      std::memcpy(buf, blob.data(), blob.size());
    }
  });
}

int mf_ls_load(mf_ls_t h, const void* buf, uint64_t size) {
  MF_TRY({
    if (!buf) throw std::runtime_error("null argument");
    ls_deserialize(get(h), static_cast<const uint8_t*>(buf), size);
  });
}

int mf_ls_apply(mf_ls_t h, void* dst, const void* src, void* stream) {
  MF_TRY({
    if (!dst || !src) throw std::runtime_error("null argument");
    if (dst == src) throw std::runtime_error("dst and src must not alias");
    ls_apply(get(h), static_cast<double*>(dst), static_cast<const double*>(src), static_cast<hipStream_t>(stream));
  });
}

int mf_ls_forward(mf_ls_t h, const void* x, int64_t n, void* y, void* stream) {
  MF_TRY({
    if (!x || !y) throw std::runtime_error("null argument");
    if (n < 0) throw std::runtime_error("negative row count");
    ls_forward(get(h), static_cast<const float*>(x), n, static_cast<float*>(y), static_cast<hipStream_t>(stream));
  });
}

int mf_ls_features(mf_ls_t h, const void* r, const void* e, void* x, void* t, void* stream) {
  MF_TRY({
    if (!r || !x || (t && !e)) throw std::runtime_error("null argument");
    ls_features(get(h), static_cast<const double*>(r), static_cast<const double*>(e), static_cast<float*>(x),
                static_cast<float*>(t), static_cast<hipStream_t>(stream));
  });
}

int mf_ls_gradient(mf_ls_t h, const void* x, const void* t, int64_t n, float* grad, double* loss, void* stream) {
  MF_TRY({
    if (!x || !t || !grad) throw std::runtime_error("null argument");
    LearnedSmoother& ls = get(h);
    hipStream_t s = static_cast<hipStream_t>(stream);
    ls_gradient(ls, static_cast<const float*>(x), static_cast<const float*>(t), n, s);
    float l;
    SMF_HIP(hipMemcpyAsync(grad, ls.d_grad.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, s));
    SMF_HIP(hipMemcpyAsync(&l, ls.d_loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
    SMF_HIP(hipStreamSynchronize(s));
    if (loss) *loss = l;
  });
}

int mf_ls_train_step(mf_ls_t h, const void* x, const void* t, int64_t n, double lr, double* loss, void* stream) {
  MF_TRY({
    if (!x || !t) throw std::runtime_error("null argument");
    if (!(lr > 0)) throw std::runtime_error("learning rate must be positive");
    LearnedSmoother& ls = get(h);
    hipStream_t s = static_cast<hipStream_t>(stream);
    ls_gradient(ls, static_cast<const float*>(x), static_cast<const float*>(t), n, s);
    ls_adam(ls, lr, s);
    float l;
    SMF_HIP(hipMemcpyAsync(&l, ls.d_loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
    SMF_HIP(hipStreamSynchronize(s));
    if (loss) *loss = l;
  });
}

int mf_ls_train(mf_ls_t h, mf_mg_t mg, const mf_train_options* opt, mf_train_stats* stats, void* stream) {
  MF_TRY({
    if (!mg) throw std::runtime_error("null argument");
    TrainOptions o;
    if (opt) {
      o.n_rhs = opt->n_rhs;
      o.pairs_per_rhs = opt->pairs_per_rhs;
      o.steps_per_pair = opt->steps_per_pair;
      o.lr = opt->lr;
      o.n_features = opt->n_features;
      o.corr_length = opt->corr_length;
      o.seed = opt->seed;
      o.loss = opt->loss;
      o.trajectory = opt->trajectory;
    }
    if (o.n_features < 1 || !(o.corr_length > 0)) throw std::runtime_error("invalid random field parameters");
    const TrainStats st = ls_train(get(h), *mg->mg, o, static_cast<hipStream_t>(stream));
    if (stats) {
      stats->pairs = st.pairs;
      stats->steps = st.steps;
      stats->first_loss = st.first_loss;
      stats->last_loss = st.last_loss;
      stats->zero_loss = st.zero_loss;
    }
  });
}

int mf_solve_fcg_learned(mf_op_t A, mf_mg_t mg, mf_ls_t h, void* x, const void* b, double rel_tol,
                         int32_t max_iter, const mf_fcg_options* opt, mf_fcg_stats* stats, void* stream) {
  MF_TRY({
    if (!A || !mg || !x || !b) throw std::runtime_error("null argument");
    require_device(A->op);
    if (A->op.kind != MF_CG_POISSON && A->op.kind != MF_DG_SIPG)
      throw std::runtime_error("solver needs a scalar CG Poisson or DG SIPG operator");
    if (mg->mg->n_fine != A->op.m()) throw std::runtime_error("multigrid does not match the operator");
    FcgOptions o;
    if (opt) {
      o.stall_ratio = opt->stall_ratio;
      o.force_fallback_at = opt->force_fallback_at;
      o.use_learned = opt->use_learned;
    }
    LearnedSmoother* ls = h ? &get(h) : nullptr;
    const FcgStats st = solve_fcg_learned(A->op, *mg->mg, ls, static_cast<double*>(x), static_cast<const double*>(b),
                                          rel_tol, max_iter, o, static_cast<hipStream_t>(stream));
    if (stats) {
      stats->base.iterations = st.base.iterations;
      stats->base.converged = st.base.converged;
      stats->base.initial_residual = st.base.r0;
      stats->base.final_residual = st.base.rn;
      stats->base.rhs_norm = st.base.b_norm;
      stats->fallback = st.fallback;
      stats->fallback_iteration = st.fallback_iteration;
      stats->learned_steps = st.learned_steps;
      stats->mean_omega = st.mean_omega;
    }
  });
}

}  // extern "C"
