// C ABI of the matrix-based comparison path (declared in include/simplexmf.h).
#include <cstring>
#include <memory>

#include "api_handles.h"
#include "matrix.h"

struct mf_sparse {
  smf::SysMatrix A;
};

using namespace smf;

extern "C" {

int mf_get_system_matrix(mf_op_t h, mf_csr_t* out) {
  MF_TRY({
    if (!h || !out) throw std::runtime_error("null argument");
    *out = nullptr;
    auto m = std::make_unique<mf_sparse>();
    assemble_cg_matrix(h->op, m->A);
    if (h->op.on_device) {
      upload_matrix(m->A);
      SMF_HIP(hipDeviceSynchronize());
    }
    *out = m.release();
  });
}

int mf_csr_vmult(mf_csr_t A, void* dst, const void* src, void* stream) {
  MF_TRY({
    if (!A) throw std::runtime_error("null argument");
    if (dst && dst == src) throw std::runtime_error("dst and src must not alias");
    if (!A->A.on_device) throw std::runtime_error("system matrix has no device data (host_only operator)");
    if (A->A.n_rows == 0) return 0;
    if (!dst || !src) throw std::runtime_error("null argument");
    launch_spmv(A->A, dst, src, static_cast<hipStream_t>(stream));
  });
}

int mf_csr_get_info(mf_csr_t A, mf_csr_info* info) {
  MF_TRY({
    if (!A || !info) throw std::runtime_error("null argument");
    std::memset(info, 0, sizeof(*info));
    info->n_rows = A->A.n_rows;
    info->nnz = A->A.nnz();
    info->precision = A->A.precision;
    info->n_row_blocks = A->A.n_blocks();
    info->block_entries = kSpmvBlockEntries;
    const double w = (double)A->A.wbytes();
    info->algorithmic_bytes = (double)info->nnz * (4 + w) + (double)info->n_rows * (8 + 2 * w);
  });
}

int mf_csr_copy_to_host(mf_csr_t A, int64_t* ptr, int32_t* col, double* val) {
  MF_TRY({
    if (!A || !ptr || (A->A.nnz() && (!col || !val))) throw std::runtime_error("null argument");
    const SysMatrix& M = A->A;
    std::memcpy(ptr, M.ptr.data(), M.ptr.size() * sizeof(int64_t));
    if (M.nnz()) {
      std::memcpy(col, M.col.data(), M.col.size() * sizeof(int32_t));
      if (M.precision == MF_F64) {
        std::memcpy(val, M.val.data(), M.val.size() * sizeof(double));
      } else {
        for (size_t e = 0; e < M.val.size(); ++e) val[e] = (double)(float)M.val[e];
      }
    }
  });
}

int mf_csr_destroy(mf_csr_t A) {
  MF_TRY({ delete A; });
}

}  // extern "C"
