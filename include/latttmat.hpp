// latttmat.hpp -- C++ host surface over the C ABI (include/latttmat.h).
//
// TransferMatrix satisfies the operator concept the reference's eigen
// driver uses (Spectra's OpType, TMat/tmat.hpp:82-120): rows(), cols(),
// perform_op(x_in, y_out), getnormfactor(), suggested_nev(), getLayerSize(),
// and get_t_instance() for the transpose (TMat/tmat.hpp:534-535).  Errors
// become latttmat::Error exceptions, so pfunc_inner's retry-on-exception
// logic (TMat/pfunc.hpp:95-102) carries over unchanged.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "latttmat.h"

namespace latttmat {

class Error : public std::runtime_error {
 public:
  Error(tm_status s, const std::string &what) : std::runtime_error(what), status_(s) {}
  tm_status status() const { return status_; }

 private:
  tm_status status_;
};

inline void check(tm_status s)
{
  if (s != TM_OK) throw Error(s, tm_last_error());
}

class TransferMatrixT;

class TransferMatrix {
 public:
  using Scalar = double;  // element type of perform_op's vectors
  using MatVecProd_T = TransferMatrixT;
  explicit TransferMatrix(const tm_params &p)
  {
    check(tm_op_create(&p, &op_));
  }
  ~TransferMatrix() { tm_op_destroy(op_); }
  TransferMatrix(const TransferMatrix &) = delete;
  TransferMatrix &operator=(const TransferMatrix &) = delete;
  TransferMatrix(TransferMatrix &&o) noexcept : op_(o.op_) { o.op_ = nullptr; }

  int64_t rows() const { return tm_op_rows(op_); }
  int64_t cols() const { return tm_op_cols(op_); }
  int suggested_nev(int neig) const { return tm_op_suggested_ncv(op_, neig); }
  double getnormfactor() const { return tm_op_normfactor(op_); }
  int getLayerSize() const { return tm_op_layer_size(op_); }
  bool is_symmetric() const { return tm_op_is_symmetric(op_) != 0; }
  // y_out = T x_in (host pointers, length cols())
  void perform_op(const double *x_in, double *y_out) const { check(tm_op_apply(op_, x_in, y_out, 0)); }
  TransferMatrixT get_t_instance() const;
  tm_op *handle() const { return op_; }

 private:
  tm_op *op_ = nullptr;
};

// The transpose instance: perform_op applies T^T (ZBTMatVecProd2D_T,
// TMat/tmat.cpp:652-662).  It borrows the operator it was made from.
class TransferMatrixT {
 public:
  using Scalar = double;
  explicit TransferMatrixT(const TransferMatrix &t) : op_(t.handle()) {}
  int64_t rows() const { return tm_op_rows(op_); }
  int64_t cols() const { return tm_op_cols(op_); }
  void perform_op(const double *x_in, double *y_out) const { check(tm_op_apply(op_, x_in, y_out, 1)); }

 private:
  tm_op *op_;
};

inline TransferMatrixT TransferMatrix::get_t_instance() const { return TransferMatrixT(*this); }

}  // namespace latttmat
