// Assembled sparse system matrix of a CG operator (the matrix-based
// comparison path) and its device product.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <vector>

#include "op.h"

namespace smf {

// entries of one row block of the product kernel (a workgroup of 256 lanes,
// eight entries per lane); a block also holds at most 256 rows
constexpr int kSpmvBlockEntries = 2048;
constexpr int kSpmvBlockRows = 256;

struct SysMatrix {
  int precision = MF_F64;
  int64_t n_rows = 0;
  std::vector<int64_t> ptr;        // n_rows + 1
  std::vector<int32_t> col;        // ascending within a row
  std::vector<double> val;         // FP64 sums (stored on the device in `precision`)
  std::vector<int32_t> blk;        // row blocks: rows [blk[b], blk[b+1])
  DevBuf d_ptr, d_col, d_val, d_blk;
  bool on_device = false;
  int64_t nnz() const { return ptr.empty() ? 0 : ptr.back(); }
  int64_t n_blocks() const { return blk.empty() ? 0 : (int64_t)blk.size() - 1; }
  size_t wbytes() const { return precision == MF_F64 ? 8 : 4; }
};

// CSR of a CG operator (scalar or components-batched, Dirichlet or Nitsche
// boundary) in the operator's DoF numbering
void assemble_cg_matrix(const Operator& op, SysMatrix& A);
void upload_matrix(SysMatrix& A);
// row blocks of the product kernel from A.ptr; false (and no blocks) when a
// row has more than kSpmvBlockEntries entries
bool make_row_blocks(SysMatrix& A);
// dst = A src, or dst += A src with add (device vectors in A.precision)
void launch_spmv(const SysMatrix& A, void* dst, const void* src, hipStream_t s, bool add = false);
// A in row blocks on the device (precision: MF_F64 or MF_F32); false when a
// row has more than kSpmvBlockEntries entries (then A is not on the device)
bool upload_blocked(const std::vector<int64_t>& ptr, const std::vector<int32_t>& col, const std::vector<double>& val,
                    int precision, SysMatrix& A);

}  // namespace smf
