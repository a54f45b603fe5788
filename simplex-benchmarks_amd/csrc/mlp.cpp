// Learned per-cell smoother: setup, training-data generation from plain
// multigrid-preconditioned CG, Adam training and flexible CG with the learned
// post-smoothing step and its automatic fallback.
#include "mlp.h"

#include <cmath>
#include <cstring>
#include <stdexcept>

namespace smf {

namespace {

double host_dot(DevBuf& partial, DevBuf& scalar, int64_t n, const double* x, const double* y, hipStream_t s) {
  blas_dot<double>(n, x, y, partial.as<double>(), scalar.as<double>(), s);
  double h;
  SMF_HIP(hipMemcpyAsync(&h, scalar.p, sizeof(double), hipMemcpyDeviceToHost, s));
  SMF_HIP(hipStreamSynchronize(s));
  return h;
}

// two dot products, one host synchronisation
void host_dot2(DevBuf& partial, DevBuf& scalar, int64_t n, const double* x0, const double* y0, const double* x1,
               const double* y1, double out[2], hipStream_t s) {
  blas_dot<double>(n, x0, y0, partial.as<double>(), scalar.as<double>(), s);
  blas_dot<double>(n, x1, y1, partial.as<double>(), scalar.as<double>() + 1, s);
  SMF_HIP(hipMemcpyAsync(out, scalar.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  SMF_HIP(hipStreamSynchronize(s));
}

void copy_d(double* dst, const double* src, int64_t n, hipStream_t s) {
  SMF_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, s));
}

uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

}  // namespace

std::unique_ptr<LearnedSmoother> make_learned_smoother(const Operator& op, uint64_t seed) {
  const bool dg = op.kind == MF_DG_SIPG;
  if (!((op.kind == MF_CG_POISSON || dg) && op.degree == 2 && op.ncomp == 1))
    throw std::runtime_error("the learned smoother is defined for the scalar P2 CG and DG operators");
  if (dg && op.dg.curved) throw std::runtime_error("learned smoother: affine cells only");
  if (op.precision != MF_F64) throw std::runtime_error("the learned smoother works on the FP64 fine operator");
  if (!op.on_device) throw std::runtime_error("operator was created host_only");
  auto ls = std::make_unique<LearnedSmoother>();
  ls->op = &op;
  const int64_t nc = dg ? op.dg.n_cells : op.cg.n_cells, nd = op.m();
  if (nc * kMlpOut >= (int64_t)1 << 31) throw std::runtime_error("too many cells for the learned smoother's int32 slots");
  ls->n_cells = nc;
  ls->n_dofs = nd;
  // cell DoFs (-1 = Dirichlet) and DoF -> slot lists in ascending slot order
  // (DG: DoFs are cell-contiguous, one slot each)
  std::vector<int32_t> cd(nc * kMlpOut);
  std::vector<int64_t> ptr(nd + 1, 0);
  for (int64_t k = 0; k < nc; ++k)
    for (int i = 0; i < kMlpOut; ++i) {
      const int64_t d = dg ? k * kMlpOut + i : op.cg.cell_dofs[k * kMlpOut + i];
      const bool free = dg || d < op.cg.n_free;
      cd[k * kMlpOut + i] = free ? (int32_t)d : -1;
      if (free) ptr[d + 1]++;
    }
  for (int64_t d = 0; d < nd; ++d) ptr[d + 1] += ptr[d];
  std::vector<int32_t> slot(ptr[nd]);
  {
    std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t k = 0; k < nc; ++k)
      for (int i = 0; i < kMlpOut; ++i)
        if (cd[k * kMlpOut + i] >= 0) slot[fill[cd[k * kMlpOut + i]]++] = (int32_t)(k * kMlpOut + i);
  }
  // geometry features: J^-T (component-major 9 x nc) and |det J| of either plan
  const double* jit = dg ? op.dg.geo.data() : op.cg.jit.data();
  auto adet = [&](int64_t k) { return dg ? op.dg.geo[(size_t)9 * nc + k] : std::fabs(op.cg.detj[k]); };
  double vol = 0;
  for (int64_t k = 0; k < nc; ++k) vol += adet(k);
  ls->h = std::cbrt(vol / (double)nc);
  std::vector<float> geo(nc * 10);
  for (int64_t k = 0; k < nc; ++k) {
    for (int e = 0; e < 9; ++e) geo[k * 10 + e] = (float)(ls->h * jit[(size_t)e * nc + k]);
    geo[k * 10 + 9] = (float)(adet(k) / (ls->h * ls->h * ls->h));
  }
  // kappa_K = mean inverse diagonal over the cell's free DoFs
  std::vector<double> diag(nd);
  {
    DevBuf dd;
    dd.alloc(nd * sizeof(double));
    launch_op(op, 1, dd.p, nullptr, nullptr);
    SMF_HIP(hipMemcpy(diag.data(), dd.p, nd * sizeof(double), hipMemcpyDeviceToHost));
  }
  std::vector<float> kappa(nc, 0.0f);
  for (int64_t k = 0; k < nc; ++k) {
    double s = 0;
    int cnt = 0;
    for (int i = 0; i < kMlpOut; ++i) {
      const int32_t d = cd[k * kMlpOut + i];
      if (d >= 0 && diag[d] != 0.0) {
        s += 1.0 / diag[d];
        cnt++;
      }
    }
    kappa[k] = cnt ? (float)(s / cnt) : 0.0f;
  }
  {
    // component-major copies for coalesced per-cell loads: [i][cell]
    std::vector<int32_t> cdt(nc * kMlpOut);
    std::vector<float> geot(nc * 10);
    for (int64_t k = 0; k < nc; ++k)
      for (int i = 0; i < kMlpOut; ++i) {
        cdt[(size_t)i * nc + k] = cd[k * kMlpOut + i];
        geot[(size_t)i * nc + k] = geo[k * 10 + i];
      }
    ls->d_cell_dofs.upload(cdt);
    ls->d_geo.upload(geot);
  }
  ls->d_slot_ptr.upload(ptr);
  ls->d_slot.upload(slot);
  ls->d_kappa.upload(kappa);
  ls->d_x.alloc(nc * kMlpIn * sizeof(float));
  ls->d_t.alloc(nc * kMlpOut * sizeof(float));
  ls->d_scale.alloc(nc * sizeof(float));
  ls->d_corr.alloc(nc * kMlpOut * sizeof(float));
  ls->d_frag.alloc(kFrags * 64 * 4 * sizeof(uint16_t));
  ls->d_dy.alloc(nc * kMlpOut * sizeof(float));
  ls->d_gpart.alloc((size_t)ls_grad_blocks() * kMlpParams * sizeof(float));
  ls->d_lpart.alloc(ls_grad_blocks() * sizeof(float));
  ls->d_grad.alloc(kMlpParams * sizeof(float));
  ls->d_loss.alloc(sizeof(float));
  ls->d_partial.alloc(blas_reduction_parts() * sizeof(double));
  ls->d_scalar.alloc(4 * sizeof(double));
  ls->d_lsc.alloc(4 * sizeof(double));
  ls->d_part2.alloc(2 * ls_dot2_parts() * sizeof(double));
  SMF_HIP(hipHostMalloc((void**)&ls->h_lsc, 4 * sizeof(double), hipHostMallocDefault));
  // He-uniform hidden layers (GELU), small output layer, zero biases
  std::vector<float> p(kMlpParams, 0.0f);
  SplitMix64 rng(seed);
  auto unif = [&]() { return (double)((rng.next() >> 11) + 0.5) * 0x1.0p-53 * 2.0 - 1.0; };
  for (int e = 0; e < kMlpH * kMlpIn; ++e) p[kOffW1 + e] = (float)(std::sqrt(6.0 / kMlpIn) * unif());
  for (int e = 0; e < kMlpH * kMlpH; ++e) p[kOffW2 + e] = (float)(std::sqrt(6.0 / kMlpH) * unif());
  for (int e = 0; e < kMlpOut * kMlpH; ++e) p[kOffW3 + e] = (float)(0.1 * std::sqrt(3.0 / kMlpH) * unif());
  ls->d_m.upload(std::vector<float>(kMlpParams, 0.0f));
  ls->d_v.upload(std::vector<float>(kMlpParams, 0.0f));
  ls_set_params(*ls, p);
  return ls;
}

void ls_set_params(LearnedSmoother& ls, const std::vector<float>& p) {
  if ((int)p.size() != kMlpParams) throw std::runtime_error("wrong number of parameters");
  if (!ls.d_params.p) ls.d_params.alloc(kMlpParams * sizeof(float));
  SMF_HIP(hipMemcpy(ls.d_params.p, p.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  ls_launch_pack(ls.d_params.as<float>(), ls.d_frag.p, nullptr);
  SMF_HIP(hipDeviceSynchronize());
}

std::vector<float> ls_get_params(const LearnedSmoother& ls) {
  std::vector<float> p(kMlpParams);
  SMF_HIP(hipDeviceSynchronize());
  SMF_HIP(hipMemcpy(p.data(), ls.d_params.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  return p;
}

std::vector<uint8_t> ls_serialize(const LearnedSmoother& ls) {
  std::vector<float> p = ls_get_params(ls), m(kMlpParams), v(kMlpParams);
  SMF_HIP(hipMemcpy(m.data(), ls.d_m.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  SMF_HIP(hipMemcpy(v.data(), ls.d_v.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  std::vector<uint8_t> out;
  auto put = [&](const void* src, size_t n) {
    const uint8_t* b = static_cast<const uint8_t*>(src);
    out.insert(out.end(), b, b + n);
  };
  put("SMFMLP01", 8);
  const uint32_t version = 1, np = kMlpParams;
  const int64_t step = ls.adam_step;
  put(&version, 4);  // the library is built for little-endian hosts only
  put(&np, 4);
  put(&step, 8);
  put(p.data(), 4 * np);
  put(m.data(), 4 * np);
  put(v.data(), 4 * np);
  for (float f : p) {
    const uint16_t b = bf16_bits(f);
    put(&b, 2);
  }
  return out;
}

void ls_deserialize(LearnedSmoother& ls, const uint8_t* data, size_t size) {
  const size_t need = 24 + (size_t)kMlpParams * 14;
  if (size != need || std::memcmp(data, "SMFMLP01", 8) != 0) throw std::runtime_error("not a learned-smoother blob");
  uint32_t version, np;
  int64_t step;
  std::memcpy(&version, data + 8, 4);
  std::memcpy(&np, data + 12, 4);
  std::memcpy(&step, data + 16, 8);
  if (version != 1 || np != (uint32_t)kMlpParams) throw std::runtime_error("unsupported learned-smoother blob version");
  std::vector<float> p(np), m(np), v(np);
  std::memcpy(p.data(), data + 24, 4 * np);
  std::memcpy(m.data(), data + 24 + 4 * np, 4 * np);
  std::memcpy(v.data(), data + 24 + 8 * np, 4 * np);
  for (uint32_t e = 0; e < np; ++e) {
    uint16_t b;
    std::memcpy(&b, data + 24 + 12 * (size_t)np + 2 * e, 2);
    if (b != bf16_bits(p[e])) throw std::runtime_error("learned-smoother blob: BF16 copy does not match the FP32 master");
  }
  ls_set_params(ls, p);
  SMF_HIP(hipMemcpy(ls.d_m.p, m.data(), 4 * np, hipMemcpyHostToDevice));
  SMF_HIP(hipMemcpy(ls.d_v.p, v.data(), 4 * np, hipMemcpyHostToDevice));
  ls.adam_step = step;
}

void ls_apply(LearnedSmoother& ls, double* corr, const double* r, hipStream_t s) {
  // features and residual norms are kept (x, scale) for a following backward pass
  // DG: every DoF lies in one cell (DoF = cell * 10 + node), so the
  // inference kernel writes the correction itself and nothing is averaged
  const bool direct = ls.op->is_dg();
  ls_launch_infer_fused(ls.n_cells, ls.d_cell_dofs.as<int32_t>(), ls.d_geo.as<float>(), ls.d_kappa.as<float>(), r,
                        nullptr, ls.d_x.as<float>(), ls.d_scale.as<float>(), ls.d_frag.p, ls.d_params.as<float>(),
                        ls.d_corr.as<float>(), s, direct ? corr : nullptr);
  if (!direct)
    ls_launch_assemble(ls.n_dofs, ls.d_slot_ptr.as<int64_t>(), ls.d_slot.as<int32_t>(), ls.d_corr.as<float>(), corr,
                       s);
}

void ls_forward(LearnedSmoother& ls, const float* x, int64_t n, float* y, hipStream_t s) {
  ls_launch_infer(n, x, ls.d_frag.p, ls.d_params.as<float>(), nullptr, nullptr, y, s);
}

void ls_features(LearnedSmoother& ls, const double* r, const double* e, float* x, float* t, hipStream_t s) {
  ls_launch_features(ls.n_cells, ls.d_cell_dofs.as<int32_t>(), ls.d_geo.as<float>(), ls.d_kappa.as<float>(), r, e,
                     x, t, ls.d_scale.as<float>(), s);
}

void ls_gradient(LearnedSmoother& ls, const float* x, const float* t, int64_t n, hipStream_t s) {
  if (n < 1) throw std::runtime_error("empty training batch");
  ls_launch_grad(n, x, t, nullptr, ls.d_params.as<float>(), ls.d_gpart.as<float>(), ls.d_lpart.as<float>(),
                 ls.d_grad.as<float>(), ls.d_loss.as<float>(), s);
}

double ls_energy_gradient(LearnedSmoother& ls, const double* rp, const double* e, hipStream_t s) {
  const Operator& A = *ls.op;
  const int64_t n = A.m(), nc = ls.n_cells;
  DevBuf c, ac;
  c.alloc(n * sizeof(double));
  ac.alloc(n * sizeof(double));
  ls_apply(ls, c.as<double>(), rp, s);  // forward as at inference (BF16 MFMA); features and scales kept
  launch_op(A, 0, ac.p, c.p, s);
  double d[2];
  host_dot2(ls.d_partial, ls.d_scalar, n, c.as<double>(), rp, c.as<double>(), ac.as<double>(), d, s);
  const double E = host_dot(ls.d_partial, ls.d_scalar, n, e, rp, s);  // |e'|_A^2 = e'.A e' = e'.r'
  const double a = d[0], q = d[1];
  if (!(E > 0)) throw std::runtime_error("training pair with zero error energy");
  if (!(q > 0) || !std::isfinite(a)) {
    // the network's correction vanishes: no direction to improve; zero gradient (Adam does not move)
    SMF_HIP(hipMemsetAsync(ls.d_grad.p, 0, kMlpParams * sizeof(float), s));
    return 1.0;
  }
  // dL/dc = -(2a / (E q)) (r' - (a/q) A c); c is the mean of the cell slots
  // of each DoF, each slot is y * s_K kappa_K
  blas_axpby<double>(n, 1.0, rp, -(a / q), ac.as<double>(), s);
  ls_launch_dy(nc, ls.d_cell_dofs.as<int32_t>(), ls.d_slot_ptr.as<int64_t>(), ls.d_scale.as<float>(),
               ls.d_kappa.as<float>(), ac.as<double>(), -2.0 * a / (E * q), ls.d_dy.as<float>(), s);
  ls_launch_grad(nc, ls.d_x.as<float>(), nullptr, ls.d_dy.as<float>(), ls.d_params.as<float>(),
                 ls.d_gpart.as<float>(), ls.d_lpart.as<float>(), ls.d_grad.as<float>(), ls.d_loss.as<float>(), s);
  SMF_HIP(hipStreamSynchronize(s));
  return 1.0 - a * a / (q * E);
}

void ls_adam(LearnedSmoother& ls, double lr, hipStream_t s) {
  ls.adam_step++;
  ls_launch_adam(ls.d_params.as<float>(), ls.d_grad.as<float>(), ls.d_m.as<float>(), ls.d_v.as<float>(), lr,
                 ls.adam_step, s);
  ls_launch_pack(ls.d_params.as<float>(), ls.d_frag.p, s);
}

TrainStats ls_train(LearnedSmoother& ls, Multigrid& mg, const TrainOptions& opt, hipStream_t s) {
  const Operator& A = *ls.op;
  if (mg.n_fine != A.m()) throw std::runtime_error("multigrid does not match the smoother's operator");
  if (opt.n_rhs < 1 || opt.pairs_per_rhs < 1 || opt.steps_per_pair < 1 || !(opt.lr > 0) ||
      (opt.loss != 0 && opt.loss != 1) || (opt.trajectory != 0 && opt.trajectory != 1))
    throw std::runtime_error("invalid training options");
  const int64_t n = A.m(), nc = ls.n_cells;
  DevBuf b, xs, x, r, z, p, q, rp, e, az, c, ac;
  for (DevBuf* v : {&b, &xs, &x, &r, &z, &p, &q, &rp, &e, &az, &c, &ac}) v->alloc(n * sizeof(double));
  TrainStats st;
  std::vector<double> bh;
  std::vector<float> th(nc * kMlpOut);
  auto read_loss = [&]() {
    float l;
    SMF_HIP(hipMemcpyAsync(&l, ls.d_loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
    SMF_HIP(hipStreamSynchronize(s));
    return (double)l;
  };
  for (int j = 0; j < opt.n_rhs; ++j) {
    assemble_rhs(A, 1, opt.seed + (uint64_t)j, opt.n_features, opt.corr_length, bh);
    SMF_HIP(hipMemcpyAsync(b.p, bh.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
    // reference solution by plain preconditioned CG (FP64 outer iteration)
    solve_cg(A, &mg, xs.as<double>(), b.as<double>(), 1e-12, 100, false, s);
    // replay CG from x = 0 (plain PCG, or flexible CG with the learned step
    // for opt.trajectory == 1); each preconditioner application gives
    // r' = r - A V(r) and e' = (x* - x) - V(r)
    const double bn = host_dot(ls.d_partial, ls.d_scalar, n, b.as<double>(), b.as<double>(), s);
    blas_fill<double>(n, 0.0, x.as<double>(), s);
    copy_d(r.as<double>(), b.as<double>(), n, s);
    double rz = 0, alpha = 0;
    for (int it = 0; it < opt.pairs_per_rhs; ++it) {
      mg_vmult(mg, z.p, r.p, s);
      launch_op(A, 0, az.p, z.p, s);
      copy_d(rp.as<double>(), r.as<double>(), n, s);
      blas_axpby<double>(n, -1.0, az.as<double>(), 1.0, rp.as<double>(), s);
      copy_d(e.as<double>(), xs.as<double>(), n, s);
      blas_axpby<double>(n, -1.0, x.as<double>(), 1.0, e.as<double>(), s);
      blas_axpby<double>(n, -1.0, z.as<double>(), 1.0, e.as<double>(), s);
      if (opt.loss == 0) {
        for (int k = 0; k < opt.steps_per_pair; ++k) {
          const double l = ls_energy_gradient(ls, rp.as<double>(), e.as<double>(), s);
          if (k == 0) {
            if (st.pairs == 0) st.first_loss = l;
            st.last_loss = l;
          }
          ls_adam(ls, opt.lr, s);
          st.steps++;
        }
        st.zero_loss = 1.0;
      } else {
        ls_features(ls, rp.as<double>(), e.as<double>(), ls.d_x.as<float>(), ls.d_t.as<float>(), s);
        for (int k = 0; k < opt.steps_per_pair; ++k) {
          ls_gradient(ls, ls.d_x.as<float>(), ls.d_t.as<float>(), nc, s);
          if (k == 0) {
            const double l = read_loss();
            if (st.pairs == 0) st.first_loss = l;
            st.last_loss = l;
          }
          ls_adam(ls, opt.lr, s);
          st.steps++;
        }
        if (j == opt.n_rhs - 1 && it == 0) {
          SMF_HIP(hipMemcpyAsync(th.data(), ls.d_t.p, th.size() * sizeof(float), hipMemcpyDeviceToHost, s));
          SMF_HIP(hipStreamSynchronize(s));
          double zl = 0;
          for (float v : th) zl += (double)v * v;
          st.zero_loss = zl / (double)nc;
        }
      }
      st.pairs++;
      if (opt.trajectory == 1) {
        // the learned step as the solver applies it: z += omega c
        ls_apply(ls, c.as<double>(), rp.as<double>(), s);
        launch_op(A, 0, ac.p, c.p, s);
        double d[2];
        host_dot2(ls.d_partial, ls.d_scalar, n, c.as<double>(), rp.as<double>(), c.as<double>(), ac.as<double>(), d,
                  s);
        const double omega = d[0] / d[1];
        if (d[1] > 0 && std::isfinite(omega)) blas_axpby<double>(n, omega, c.as<double>(), 1.0, z.as<double>(), s);
      }
      // (F)CG step
      const double rz_new = host_dot(ls.d_partial, ls.d_scalar, n, r.as<double>(), z.as<double>(), s);
      if (it == 0) {
        copy_d(p.as<double>(), z.as<double>(), n, s);
      } else {
        const double beta = opt.trajectory == 1
                                ? -alpha * host_dot(ls.d_partial, ls.d_scalar, n, z.as<double>(), q.as<double>(), s) / rz
                                : rz_new / rz;
        blas_axpby<double>(n, 1.0, z.as<double>(), beta, p.as<double>(), s);
      }
      rz = rz_new;
      launch_op(A, 0, q.p, p.p, s);
      const double pq = host_dot(ls.d_partial, ls.d_scalar, n, p.as<double>(), q.as<double>(), s);
      alpha = rz / pq;
      blas_axpby<double>(n, alpha, p.as<double>(), 1.0, x.as<double>(), s);
      blas_axpby<double>(n, -alpha, q.as<double>(), 1.0, r.as<double>(), s);
      const double rn = host_dot(ls.d_partial, ls.d_scalar, n, r.as<double>(), r.as<double>(), s);
      if (!(rn > 1e-22 * bn)) break;
    }
  }
  SMF_HIP(hipStreamSynchronize(s));
  return st;
}

FcgStats solve_fcg_learned(const Operator& A, Multigrid& mg, LearnedSmoother* ls, double* x, const double* b,
                           double rel_tol, int max_iter, const FcgOptions& fo, hipStream_t s) {
  if (A.precision != MF_F64) throw std::runtime_error("the outer solver runs in FP64");
  if (ls && ls->op->m() != A.m()) throw std::runtime_error("learned smoother does not match the operator");
  const int64_t n = A.m();
  DevBuf r, z, p, q, az, c, ac, partial, scalar;
  for (DevBuf* v : {&r, &z, &p, &q, &az, &c, &ac}) v->alloc(n * sizeof(double));
  partial.alloc(blas_reduction_parts() * sizeof(double));
  scalar.alloc(2 * sizeof(double));
  FcgStats st;
  bool learned = ls != nullptr && fo.use_learned != 0;
  bool pending = false;  // a learned step whose omega has not been read back yet
  double omega_sum = 0;
  double* lsc = ls ? ls->d_lsc.as<double>() : nullptr;
  const double* lsh = ls ? ls->h_lsc : nullptr;
  auto fall_back = [&](int it) {
    if (!learned) return;
    learned = false;
    st.fallback = 1;
    st.fallback_iteration = it;
  };
  // read back the last learned step's omega with a synchronisation the
  // iteration makes anyway; an unusable step switches the network off
  auto collect = [&](int it) {
    if (!pending) return;
    pending = false;
    if (lsh[3] != 0.0) {
      fall_back(it);
    } else {
      st.learned_steps++;
      omega_sum += lsh[2];
    }
  };
  // z = V(r) [+ omega L(r - A V(r))]; returns true when az = A z was formed
  auto precond = [&]() -> bool {
    mg_vmult(mg, z.p, r.p, s);
    if (!learned) return false;
    launch_op(A, 0, az.p, z.p, s);
    // c = L(r - az): features formed inside the inference kernel
    const bool direct = A.is_dg();  // one cell per DoF: no averaging
    ls_launch_infer_fused(ls->n_cells, ls->d_cell_dofs.as<int32_t>(), ls->d_geo.as<float>(),
                          ls->d_kappa.as<float>(), r.as<double>(), az.as<double>(), nullptr, nullptr, ls->d_frag.p,
                          ls->d_params.as<float>(), ls->d_corr.as<float>(), s, direct ? c.as<double>() : nullptr);
    if (!direct)
      ls_launch_assemble(ls->n_dofs, ls->d_slot_ptr.as<int64_t>(), ls->d_slot.as<int32_t>(), ls->d_corr.as<float>(),
                         c.as<double>(), s);
    launch_op(A, 0, ac.p, c.p, s);
    ls_launch_omega(n, ls->d_part2.as<double>(), lsc, c.as<double>(), r.as<double>(), az.as<double>(),
                    ac.as<double>(), z.as<double>(), az.as<double>(), s);
    SMF_HIP(hipMemcpyAsync(ls->h_lsc, lsc, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    pending = true;
    return true;
  };
  blas_fill<double>(n, 0.0, x, s);
  copy_d(r.as<double>(), b, n, s);
  st.base.b_norm = std::sqrt(host_dot(partial, scalar, n, b, b, s));
  st.base.r0 = st.base.rn = st.base.b_norm;
  const double tol = rel_tol * st.base.b_norm;
  if (st.base.r0 <= tol) {
    st.base.converged = 1;
    return st;
  }
  if (fo.force_fallback_at == 0) fall_back(0);
  // alpha, beta and the dot products stay in device memory (S); the host
  // reads |r|^2 once per iteration (the stopping test and stall detector)
  // and, with it, the last learned step's omega
  DevBuf S;
  S.alloc(kS_N * sizeof(double));
  double* Sd = S.as<double>();
  bool have_az = precond();
  copy_d(p.as<double>(), z.as<double>(), n, s);
  if (have_az) copy_d(q.as<double>(), az.as<double>(), n, s);
  else launch_op(A, 0, q.p, p.p, s);
  blas_dot<double>(n, r.as<double>(), z.as<double>(), partial.as<double>(), Sd + kS_RZ, s);
  double rprev = st.base.r0;
  int slow = 0;
  for (int it = 1; it <= max_iter; ++it) {
    blas_dot<double>(n, p.as<double>(), q.as<double>(), partial.as<double>(), Sd + kS_PQ, s);
    blas_krylov_scalars(Sd, 0, s);                                                 // alpha = rz / pq
    blas_axpby_dev(n, Sd, kS_ALPHA, 1.0, p.as<double>(), -1, x, s);               // x += alpha p
    blas_axpby_dev(n, Sd, kS_ALPHA, -1.0, q.as<double>(), -1, r.as<double>(), s);  // r -= alpha q
    st.base.rn = std::sqrt(host_dot(partial, scalar, n, r.as<double>(), r.as<double>(), s));
    collect(it - 1);  // the learned step of the previous preconditioner application
    st.base.iterations = it;
    if (st.base.rn <= tol || !std::isfinite(st.base.rn)) {
      st.base.converged = st.base.rn <= tol;
      break;
    }
    // stall detector: two consecutive weak reductions switch the learned step off
    slow = st.base.rn > fo.stall_ratio * rprev ? slow + 1 : 0;
    if (slow >= 2 || it == fo.force_fallback_at) fall_back(it);
    rprev = st.base.rn;
    have_az = precond();
    blas_dot<double>(n, r.as<double>(), z.as<double>(), partial.as<double>(), Sd + kS_RZN, s);
    blas_dot<double>(n, z.as<double>(), q.as<double>(), partial.as<double>(), Sd + kS_ZQ, s);
    // flexible CG (Notay 2000): beta = -alpha z.q / rz
    blas_krylov_scalars(Sd, 2, s);
    blas_axpby_dev(n, Sd, -1, 1.0, z.as<double>(), kS_BETA, p.as<double>(), s);  // p = z + beta p
    if (have_az) blas_axpby_dev(n, Sd, -1, 1.0, az.as<double>(), kS_BETA, q.as<double>(), s);
    else launch_op(A, 0, q.p, p.p, s);
  }
  SMF_HIP(hipStreamSynchronize(s));
  collect(st.base.iterations);
  st.mean_omega = st.learned_steps ? omega_sum / st.learned_steps : 0.0;
  return st;
}

}  // namespace smf
