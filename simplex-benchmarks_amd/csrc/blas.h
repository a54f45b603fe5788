// Launchers of the vector / transfer / reduction kernels (blas_kernels.hip).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace smf {

template <typename T>
void blas_axpby(int64_t n, double a, const T* x, double b, T* y, hipStream_t s);  // y = a x + b y (FP64 math)
template <typename T>
void blas_axpby_t(int64_t n, T a, const T* x, T b, T* y, hipStream_t s);         // y = a x + b y (T math)
template <typename TO, typename TI>
void blas_convert(int64_t n, const TI* x, TO* y, hipStream_t s);
template <typename T>
void blas_scale(int64_t n, const T* d, const T* x, T* y, hipStream_t s);         // y = d .* x
template <typename T>
void blas_fill(int64_t n, T v, T* y, hipStream_t s);
// *out = x . y in FP64, deterministic; partial holds blas_reduction_parts() doubles
template <typename T>
void blas_dot(int64_t n, const T* x, const T* y, double* partial, double* out, hipStream_t s);
int blas_reduction_parts();
template <typename T>
void blas_csr(int64_t nrows, const int64_t* ptr, const int32_t* col, const T* val, const T* x, T* y, bool add,
              hipStream_t s);
template <typename T>
void blas_cheby_step(int64_t n, T c1, T c2, const T* dinv, const T* b, const T* ax, T* d, T* x, bool first,
                     hipStream_t s);
template <typename T>
void blas_cheby_zero(int64_t n, T c2, const T* dinv, const T* b, T* d, T* x, hipStream_t s);  // d = x = c2 Dinv b
// y[0..nf) = M x[0..nf) for a dense row-major nf x nf matrix, y[nf..n) = x[nf..n)
void blas_dense_gemv(int nf, int n, const double* M, const double* x, double* y, hipStream_t s);

// Krylov scalars kept in device memory, so that a (flexible) CG iteration
// synchronises with the host once (for its residual norm).  Slots of the
// scalar block S:
enum KrylovSlot { kS_RZ = 0, kS_PQ = 1, kS_RR = 2, kS_RZN = 3, kS_ZQ = 4, kS_ALPHA = 5, kS_BETA = 6, kS_N = 8 };
// y = (sa * S[ia]) x + S[ib] y in FP64; an index < 0 reads as 1
void blas_axpby_dev(int64_t n, const double* S, int ia, double sa, const double* x, int ib, double* y, hipStream_t s);
// mode 0: S[ALPHA] = S[RZ] / S[PQ];  mode 1: S[BETA] = S[RZN] / S[RZ], S[RZ] = S[RZN];
// mode 2 (flexible, Notay): S[BETA] = -S[ALPHA] S[ZQ] / S[RZ], S[RZ] = S[RZN]
void blas_krylov_scalars(double* S, int mode, hipStream_t s);

}  // namespace smf
