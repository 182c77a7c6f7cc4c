// Host side of the persistent H-LLt solve sweeps: the item lists of the
// forward and backward sweeps (built once per factorisation from the block
// tree), the inverses of the diagonal leaves, and the launches.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "engine.hpp"
#include "sweep.hpp"

namespace hmx {

class SweepSolver {
 public:
  // Plan for the LLt factor of M, or nullptr (with the reason) when the
  // partition has a shape the sweeps do not handle (leaf clusters above 256
  // rows, Rk capacities above 256, full leaves spanning several clusters).
  static std::unique_ptr<SweepSolver> build(const Matrix& M, std::string* why);

  // x <- (L L^T)^{-1} x in place, x N x nrhs with leading dimension ldx
  // (device, hmat numbering), on stream s.  The inverses of the diagonal
  // leaves are formed on the first call for the store's precision.
  template <typename T>
  void solve(Matrix& M, const Store<T>& st, T* x, int nrhs, int ldx, hipStream_t s);

  // Raises if a sweep gave up waiting (synchronises the stream).
  void check(hipStream_t s);

  int rows() const { return nrows_; }

 private:
  struct Dir {
    DevBuf items, rows, contrib, chunks, reds;
    int nitems = 0;
  };
  Dir dir_[2];  // 0 forward, 1 backward
  DevBuf diag_rows_;       // SwRow per cluster in hmat order (for the inverses)
  std::vector<int> diag_leaf_;
  std::vector<SwRow> hdiag_;
  int nrows_ = 0, nslots_ = 0;
  int64_t wait_ticks_ = 0;
  int64_t z_elems_ = 0, zpart_elems_ = 0, inv_elems_ = 0;
  DevBuf z_, zpart_, sync_, fail_, flags_;
  DevBuf inv64_, inv32_;
  DevBuf trsm_tasks_;
  bool inv_ready_[2] = {false, false};
  template <typename T>
  void ensure_inverse(const Store<T>& st, hipStream_t s);
};

}  // namespace hmx
