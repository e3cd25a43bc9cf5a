// PyTorch-ROCm extension: the reference's pybind surface on top of the C ABI.
//
// Mirrors PYBIND11_MODULE in iterative/BiCGSTAB/spmv.cpp:892-905
// (csr_tensor(keys, values, row_ptr, col_ind, out_vals, num_rows, num_cols) with
// .mv(x, b), .cgd(b, x), .bicgstab(b, x)), the Richardson module's iter_solve
// (iterative/richardson/spmv.cpp:499-512) and the standalone converter
// coo_csr.coo_csr(...) (spmv_kernel/coo_csr/coo_csr/coo_csr/coo_to_csr.cpp:282-283).
// Arguments keep the reference's order and in-place meaning (the last tensor
// argument is the output). Tensors may live on the CPU (copied) or on the
// current ROCm device (used in place); work runs on torch's current stream.
#include <torch/extension.h>
#include <c10/hip/HIPStream.h>
#include <c10/hip/HIPGuard.h>

#include <string>
#include <vector>

#include "../../include/mpsparse_amd.h"

namespace py = pybind11;

namespace {

void check(msp_status s, const char* what) {
  TORCH_CHECK(s == MSP_OK, what, " failed (status ", (int)s, "): ", msp_last_error());
}

void* stream_for(int device) {
  return (void*)c10::hip::getCurrentHIPStream((c10::DeviceIndex)device).stream();
}

int pick_device(std::initializer_list<const torch::Tensor*> ts) {
  int dev = -1;
  for (const torch::Tensor* t : ts) {
    if (t->defined() && t->is_cuda()) {
      TORCH_CHECK(dev < 0 || dev == t->get_device(), "tensors on different GPUs");
      dev = t->get_device();
    }
  }
  if (dev < 0) dev = (int)c10::hip::current_device();
  return dev;
}

void expect(const torch::Tensor& t, c10::ScalarType dt, int64_t numel, const char* name) {
  TORCH_CHECK(t.scalar_type() == dt, name, ": expected dtype ", c10::toString(dt), ", got ",
              c10::toString(t.scalar_type()));
  TORCH_CHECK(t.is_contiguous(), name, ": must be contiguous");
  TORCH_CHECK(numel < 0 || t.numel() == numel, name, ": expected ", numel, " elements, got ", t.numel());
  TORCH_CHECK(t.is_cpu() || t.is_cuda(), name, ": must be a CPU or ROCm tensor");
}

int32_t value_type_of(const std::string& v) {
  if (v == "f32" || v == "fp32" || v == "float32") return MSP_VALUES_F32;
  if (v == "bf16" || v == "bfloat16") return MSP_VALUES_BF16;
  if (v == "f16" || v == "fp16" || v == "float16") return MSP_VALUES_F16;
  TORCH_CHECK(false, "unknown value type '", v, "' (f32, bf16, f16)");
  return 0;
}

py::dict info_dict(const msp_solve_info& in) {
  py::dict d;
  d["iterations"] = in.iterations;
  d["status"] = in.status == MSP_SOLVE_CONVERGED   ? "converged"
                : in.status == MSP_SOLVE_BREAKDOWN ? "breakdown"
                : in.status == MSP_SOLVE_STAGNATED ? "stagnated"
                                                   : "maxiter";
  d["verifications"] = in.verifications;
  d["relres"] = in.relres;
  d["rr"] = in.rr;
  d["bb"] = in.bb;
  d["seconds"] = in.seconds;
  return d;
}

py::dict refine_dict(const msp_refine_info& info) {
  py::dict d;
  d["outer_iterations"] = info.outer_iterations;
  d["inner_iterations"] = info.inner_iterations;
  d["status"] = info.status == MSP_SOLVE_CONVERGED   ? "converged"
                : info.status == MSP_SOLVE_BREAKDOWN ? "breakdown"
                : info.status == MSP_SOLVE_STAGNATED ? "stagnated"
                                                     : "maxiter";
  d["fallbacks"] = info.fallbacks;
  d["relres"] = info.relres;
  d["seconds"] = info.seconds;
  return d;
}

class CsrTensor {
 public:
  CsrTensor(torch::Tensor keys, torch::Tensor values, torch::Tensor row_ptr, torch::Tensor col_ind,
            torch::Tensor out_vals, int64_t num_rows, int64_t num_cols) {
    TORCH_CHECK(num_rows >= 0 && num_cols >= 0 && num_rows < (1ll << 32) && num_cols < (1ll << 32),
                "num_rows / num_cols out of range");
    expect(keys, torch::kInt64, -1, "keys");
    expect(values, torch::kFloat32, keys.numel(), "values");
    device_ = pick_device({&keys, &values});
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    void* s = stream_for(device_);
    {
      py::gil_scoped_release nogil;
      check(msp_csr_create_coo(&h_, reinterpret_cast<const uint64_t*>(keys.data_ptr<int64_t>()),
                               values.data_ptr<float>(), (uint64_t)keys.numel(), (uint32_t)num_rows,
                               (uint32_t)num_cols, s),
            "csr_tensor");
    }
    // Unlike the reference (which ignores these buffers), write the CSR out
    // when the caller passes buffers of the right size.
    export_into(row_ptr, col_ind, out_vals);
  }

  ~CsrTensor() {
    if (h_) msp_csr_destroy(h_);
  }

  void export_into(torch::Tensor row_ptr, torch::Tensor col_ind, torch::Tensor out_vals) {
    uint32_t n = 0, m = 0;
    uint64_t nnz = 0;
    check(msp_csr_shape(h_, &n, &m, &nnz), "shape");
    uint32_t* rp = nullptr;
    uint32_t* ci = nullptr;
    float* ov = nullptr;
    if (row_ptr.defined() && row_ptr.numel() == (int64_t)n + 1) {
      expect(row_ptr, torch::kInt32, (int64_t)n + 1, "row_ptr");
      rp = reinterpret_cast<uint32_t*>(row_ptr.data_ptr<int32_t>());
    }
    if (col_ind.defined() && col_ind.numel() == (int64_t)nnz && nnz > 0) {
      expect(col_ind, torch::kInt32, (int64_t)nnz, "col_ind");
      ci = reinterpret_cast<uint32_t*>(col_ind.data_ptr<int32_t>());
    }
    if (out_vals.defined() && out_vals.numel() == (int64_t)nnz && nnz > 0) {
      expect(out_vals, torch::kFloat32, (int64_t)nnz, "out_vals");
      ov = out_vals.data_ptr<float>();
    }
    if (!rp && !ci && !ov) return;
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    void* s = stream_for(device_);
    py::gil_scoped_release nogil;
    check(msp_csr_export(h_, rp, ci, ov, s), "export");
  }

  void mv(torch::Tensor x, torch::Tensor b, const std::string& values) {
    uint32_t n = 0, m = 0;
    check(msp_csr_shape(h_, &n, &m, nullptr), "shape");
    expect(x, torch::kFloat32, m, "x");
    expect(b, torch::kFloat32, n, "b");
    int32_t vt = value_type_of(values);
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    void* s = stream_for(device_);
    py::gil_scoped_release nogil;
    if (vt != MSP_VALUES_F32) check(msp_csr_prepare_values(h_, vt, s), "prepare_values");
    check(msp_csr_spmv(h_, x.data_ptr<float>(), b.data_ptr<float>(), vt, s), "mv");
  }

  msp_solve_opts make_opts(double tol, int64_t maxiter, int64_t check_every, const std::string& criterion,
                           bool use_x0, const std::string& precond, const std::string& values, bool verify,
                           const std::string& cg_form = "auto") {
    msp_solve_opts o;
    msp_default_opts(&o);
    o.tol = tol;
    o.maxiter = (int32_t)maxiter;
    o.check_every = (int32_t)check_every;
    if (criterion == "relres") o.criterion = MSP_CRIT_RELRES;
    else if (criterion == "reference") o.criterion = MSP_CRIT_REFERENCE;
    else TORCH_CHECK(false, "criterion must be 'relres' or 'reference'");
    o.use_x0 = use_x0 ? 1 : 0;
    if (precond == "none") o.precond = MSP_PRECOND_NONE;
    else if (precond == "jacobi") o.precond = MSP_PRECOND_JACOBI;
    else TORCH_CHECK(false, "precond must be 'none' or 'jacobi'");
    o.values = value_type_of(values);
    o.verify = verify ? 1 : 0;
    if (cg_form == "auto") o.cg_form = 0;
    else if (cg_form == "fused") o.cg_form = 1;
    else if (cg_form == "separate") o.cg_form = 2;
    else TORCH_CHECK(false, "cg_form must be 'auto', 'fused' or 'separate'");
    return o;
  }

  py::dict solve(bool bicg, torch::Tensor b, torch::Tensor x, double tol, int64_t maxiter, int64_t check_every,
                 const std::string& criterion, bool use_x0, const std::string& precond, const std::string& values,
                 bool verify, const std::string& cg_form = "auto") {
    uint32_t n = 0, m = 0;
    check(msp_csr_shape(h_, &n, &m, nullptr), "shape");
    expect(b, torch::kFloat32, n, "b");
    expect(x, torch::kFloat32, m, "x");
    msp_solve_opts o = make_opts(tol, maxiter, check_every, criterion, use_x0, precond, values, verify, cg_form);
    msp_solve_info info{};
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    void* s = stream_for(device_);
    {
      py::gil_scoped_release nogil;
      if (o.values != MSP_VALUES_F32) check(msp_csr_prepare_values(h_, o.values, s), "prepare_values");
      if (bicg) check(msp_csr_bicgstab(h_, b.data_ptr<float>(), x.data_ptr<float>(), &o, &info, s), "bicgstab");
      else check(msp_csr_cg(h_, b.data_ptr<float>(), x.data_ptr<float>(), &o, &info, s), "cgd");
    }
    return info_dict(info);
  }

  void richardson(torch::Tensor b, torch::Tensor x, double omega, int64_t iters, bool use_x0,
                  const std::string& values) {
    uint32_t n = 0, m = 0;
    check(msp_csr_shape(h_, &n, &m, nullptr), "shape");
    expect(b, torch::kFloat32, n, "b");
    expect(x, torch::kFloat32, m, "x");
    TORCH_CHECK(iters >= 0 && iters < (1ll << 31), "iters out of range");
    int32_t vt = value_type_of(values);
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    void* s = stream_for(device_);
    py::gil_scoped_release nogil;
    if (vt != MSP_VALUES_F32) check(msp_csr_prepare_values(h_, vt, s), "prepare_values");
    check(msp_csr_richardson(h_, b.data_ptr<float>(), x.data_ptr<float>(), (float)omega, (int32_t)iters,
                             use_x0 ? 1 : 0, vt, s),
          "richardson");
  }

  py::dict solve_refined(torch::Tensor b, torch::Tensor x, double tol, double inner_tol, int64_t max_outer,
                         int64_t inner_maxiter, const std::string& solver, const std::string& values,
                         const std::string& precond, bool use_x0) {
    uint32_t n = 0, m = 0;
    check(msp_csr_shape(h_, &n, &m, nullptr), "shape");
    expect(b, torch::kFloat64, n, "b");
    expect(x, torch::kFloat64, m, "x");
    msp_refine_opts o;
    msp_default_refine_opts(&o);
    o.tol = tol;
    o.inner_tol = inner_tol;
    o.max_outer = (int32_t)max_outer;
    o.inner_maxiter = (int32_t)inner_maxiter;
    TORCH_CHECK(solver == "cg" || solver == "bicgstab", "solver must be 'cg' or 'bicgstab'");
    o.inner_solver = solver == "cg" ? 0 : 1;
    o.values = value_type_of(values);
    TORCH_CHECK(precond == "none" || precond == "jacobi", "precond must be 'none' or 'jacobi'");
    o.precond = precond == "jacobi" ? MSP_PRECOND_JACOBI : MSP_PRECOND_NONE;
    o.use_x0 = use_x0 ? 1 : 0;
    msp_refine_info info{};
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    void* s = stream_for(device_);
    {
      py::gil_scoped_release nogil;
      check(msp_csr_solve_refined(h_, b.data_ptr<double>(), x.data_ptr<double>(), &o, &info, s), "solve_refined");
    }
    return refine_dict(info);
  }

  void set_format(const std::string& f) {
    TORCH_CHECK(f == "auto" || f == "csr", "format must be 'auto' or 'csr'");
    check(msp_csr_set_format(h_, f == "auto" ? MSP_FORMAT_AUTO : MSP_FORMAT_CSR), "set_format");
  }
  std::string format() const {
    int32_t f = 0;
    check(msp_csr_get_format(h_, &f), "format");
    return f == 3 ? "dia_symmetric" : (f == 2 ? "dia" : "csr");
  }
  int64_t num_rows() const {
    uint32_t n = 0;
    msp_csr_shape(h_, &n, nullptr, nullptr);
    return n;
  }
  int64_t num_cols() const {
    uint32_t m = 0;
    msp_csr_shape(h_, nullptr, &m, nullptr);
    return m;
  }
  int64_t nnz() const {
    uint64_t z = 0;
    msp_csr_shape(h_, nullptr, nullptr, &z);
    return (int64_t)z;
  }
  int device() const { return device_; }
  uintptr_t handle() const { return reinterpret_cast<uintptr_t>(h_); }
  msp_csr raw() const { return h_; }

 private:
  msp_csr h_ = nullptr;
  int device_ = 0;
};

void* stream_of(int device) { return stream_for(device); }

// Learned / classical FSAI preconditioner bound to one csr_tensor (kept alive
// by a reference held on the Python side).
class FsaiTensor {
 public:
  explicit FsaiTensor(CsrTensor& A) : A_(A.raw()), device_(A.device()) {
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    void* s = stream_of(device_);
    py::gil_scoped_release nogil;
    check(msp_fsai_create(&h_, A_, s), "fsai");
    check(msp_fsai_num_edges(h_, &ne_), "fsai edges");
  }
  ~FsaiTensor() {
    if (h_) msp_fsai_destroy(h_);
  }
  int64_t num_edges() const { return (int64_t)ne_; }
  torch::TensorOptions dopt(c10::ScalarType t) const {
    return torch::TensorOptions().device(torch::kCUDA, device_).dtype(t);
  }
  py::tuple edges() {
    auto r = torch::empty({(int64_t)ne_}, dopt(torch::kInt32));
    auto cc = torch::empty({(int64_t)ne_}, dopt(torch::kInt32));
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    check(msp_fsai_edges(h_, reinterpret_cast<uint32_t*>(r.data_ptr<int32_t>()),
                         reinterpret_cast<uint32_t*>(cc.data_ptr<int32_t>()), stream_of(device_)),
          "edges");
    return py::make_tuple(r, cc);
  }
  py::tuple features() {
    auto f = torch::empty({(int64_t)ne_, 8}, dopt(torch::kFloat32));
    auto sc = torch::empty({(int64_t)ne_}, dopt(torch::kFloat32));
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    check(msp_fsai_features(h_, f.data_ptr<float>(), sc.data_ptr<float>(), stream_of(device_)), "features");
    return py::make_tuple(f, sc);
  }
  int classical() {
    int32_t fb = 0;
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    check(msp_fsai_classical(h_, &fb, stream_of(device_)), "classical");
    return fb;
  }
  int set_values(torch::Tensor g) {
    expect(g, torch::kFloat32, (int64_t)ne_, "g");
    int32_t np = 0;
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    check(msp_fsai_set_values(h_, g.data_ptr<float>(), &np, stream_of(device_)), "set_values");
    return np;
  }
  // "coupled8" or "stated5" (msp_fsai_features)
  void select_features(const std::string& name) {
    int32_t set = -1;
    if (name == "coupled8") set = MSP_FSAI_FEATURES_COUPLED8;
    else if (name == "stated5") set = MSP_FSAI_FEATURES_STATED5;
    TORCH_CHECK(set >= 0, "feature set must be 'coupled8' or 'stated5', got '", name, "'");
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    check(msp_fsai_select_features(h_, set, stream_of(device_)), "select_features");
  }
  std::string feature_set() const {
    int32_t set = 0;
    check(msp_fsai_get_features(h_, &set), "feature_set");
    return set == MSP_FSAI_FEATURES_STATED5 ? "stated5" : "coupled8";
  }
  // G from an edge-network output o (g = o / sqrt(a_jj), then the row
  // normalisation of the stated-5 feature set); returns the non-positive count
  int set_network_output(torch::Tensor out) {
    expect(out, torch::kFloat32, (int64_t)ne_, "out");
    int32_t np = 0;
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    check(msp_fsai_set_network_output(h_, out.data_ptr<float>(), &np, stream_of(device_)), "set_network_output");
    return np;
  }
  int set_network(torch::Tensor params) {
    expect(params, torch::kFloat32, msp_mlp_num_params(), "params");
    TORCH_CHECK(params.is_cuda(), "params must be a ROCm tensor");
    int32_t np = 0;
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    check(msp_fsai_apply_network(h_, params.data_ptr<float>(), &np, stream_of(device_)), "set_network");
    return np;
  }
  torch::Tensor values() {
    auto g = torch::empty({(int64_t)ne_}, dopt(torch::kFloat32));
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    check(msp_fsai_export_values(h_, g.data_ptr<float>(), stream_of(device_)), "values");
    return g;
  }
  void apply(torch::Tensor r, torch::Tensor z) {
    expect(r, torch::kFloat32, -1, "r");
    expect(z, torch::kFloat32, r.numel(), "z");
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    check(msp_fsai_apply(h_, r.data_ptr<float>(), z.data_ptr<float>(), stream_of(device_)), "apply");
  }
  double probe_loss_grad(torch::Tensor W, torch::Tensor grad) {
    expect(W, torch::kFloat32, -1, "W");
    TORCH_CHECK(W.dim() == 2, "W must be (k, n)");
    expect(grad, torch::kFloat32, (int64_t)ne_, "grad");
    TORCH_CHECK(grad.is_cuda(), "grad must be a ROCm tensor");
    double loss = 0.0;
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    void* s = stream_of(device_);
    py::gil_scoped_release nogil;
    check(msp_fsai_probe_loss_grad(h_, W.data_ptr<float>(), (int32_t)W.size(0), grad.data_ptr<float>(), &loss, s),
          "probe_loss_grad");
    return loss;
  }
  double train_step(torch::Tensor params, torch::Tensor W, torch::Tensor grad_params) {
    expect(params, torch::kFloat32, msp_mlp_num_params(), "params");
    expect(grad_params, torch::kFloat32, msp_mlp_num_params(), "grad_params");
    expect(W, torch::kFloat32, -1, "W");
    TORCH_CHECK(W.dim() == 2, "W must be (k, n)");
    double loss = 0.0;
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    void* s = stream_of(device_);
    py::gil_scoped_release nogil;
    check(msp_fsai_train_step(h_, params.data_ptr<float>(), W.data_ptr<float>(), (int32_t)W.size(0),
                              grad_params.data_ptr<float>(), &loss, s),
          "train_step");
    return loss;
  }
  msp_fsai raw() const { return h_; }
  msp_csr matrix() const { return A_; }

 private:
  msp_csr A_ = nullptr;
  int device_ = 0;
  msp_fsai h_ = nullptr;
  uint64_t ne_ = 0;
};

py::dict pcg_fsai(CsrTensor& A, FsaiTensor& F, torch::Tensor b, torch::Tensor x, double tol, int64_t maxiter,
                  int64_t check_every, bool use_x0, const std::string& values, bool verify) {
  TORCH_CHECK(F.matrix() == A.raw(), "the FSAI preconditioner was built for another matrix");
  const int64_t n = A.num_rows();
  expect(b, torch::kFloat32, n, "b");
  expect(x, torch::kFloat32, n, "x");
  msp_solve_opts o;
  msp_default_opts(&o);
  o.tol = tol;
  o.maxiter = (int32_t)maxiter;
  o.check_every = (int32_t)check_every;
  o.use_x0 = use_x0 ? 1 : 0;
  o.values = value_type_of(values);
  o.verify = verify ? 1 : 0;
  msp_solve_info info{};
  c10::hip::HIPGuard guard((c10::DeviceIndex)A.device());
  void* s = stream_of(A.device());
  {
    py::gil_scoped_release nogil;
    if (o.values != MSP_VALUES_F32) check(msp_csr_prepare_values(A.raw(), o.values, s), "prepare_values");
    check(msp_csr_pcg_fsai(A.raw(), F.raw(), b.data_ptr<float>(), x.data_ptr<float>(), &o, &info, s), "pcg_fsai");
  }
  py::dict d = info_dict(info);
  d["fell_back_to_cg"] = info.reserved == 1;
  return d;
}

py::dict pcg_fsai_refined(CsrTensor& A, FsaiTensor& F, torch::Tensor b, torch::Tensor x, double tol, double inner_tol,
                          int64_t max_outer, int64_t inner_maxiter, const std::string& values, bool use_x0) {
  TORCH_CHECK(F.matrix() == A.raw(), "the FSAI preconditioner was built for another matrix");
  const int64_t n = A.num_rows();
  expect(b, torch::kFloat64, n, "b");
  expect(x, torch::kFloat64, n, "x");
  msp_refine_opts o;
  msp_default_refine_opts(&o);
  o.tol = tol;
  o.inner_tol = inner_tol;
  o.max_outer = (int32_t)max_outer;
  o.inner_maxiter = (int32_t)inner_maxiter;
  o.values = value_type_of(values);
  o.use_x0 = use_x0 ? 1 : 0;
  msp_refine_info info{};
  c10::hip::HIPGuard guard((c10::DeviceIndex)A.device());
  void* s = stream_of(A.device());
  {
    py::gil_scoped_release nogil;
    check(msp_csr_solve_refined_fsai(A.raw(), F.raw(), b.data_ptr<double>(), x.data_ptr<double>(), &o, &info, s),
          "pcg_fsai_refined");
  }
  return refine_dict(info);
}

// Batched Jacobi-PCG over the diagonal blocks of a csr_tensor (kept alive by a
// reference held on the Python side).
// Drop-in for SparseQRSolver (sparse_bridge/bridge.cpp:14-191): the
// constructor orders and factors the CSC matrix (int64 indices, float64
// values; CPU or ROCm tensors) on the GPU; solve(B) takes a vector or a matrix
// of right-hand sides in any layout and returns X on B's device (num_cols rows,
// column-major for a matrix, as bridge.cpp:128-152 returns it);
// get_grad_a_helper(col_starts, row_indices, grad_b, x, n_rows, n_cols)
// returns -grad_b[i] * x[j] per CSC entry (i, j) (bridge.cpp:160-189).
class QrSolver {
 public:
  QrSolver(torch::Tensor col_starts, torch::Tensor row_indices, torch::Tensor values, int64_t n_rows,
           int64_t n_cols)
      : m_(n_rows), n_(n_cols) {
    TORCH_CHECK(n_rows > 0 && n_cols > 0, "SparseQRSolver: empty matrix");
    expect(col_starts, torch::kInt64, n_cols + 1, "col_starts");
    expect(row_indices, torch::kInt64, -1, "row_indices");
    expect(values, torch::kFloat64, row_indices.numel(), "values");
    device_ = pick_device({&col_starts, &row_indices, &values});
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    void* s = stream_of(device_);
    py::gil_scoped_release nogil;
    check(msp_qr_create_csc(&h_, col_starts.data_ptr<int64_t>(), row_indices.data_ptr<int64_t>(),
                            values.data_ptr<double>(), n_rows, n_cols, s),
          "SparseQRSolver");
  }
  ~QrSolver() {
    if (h_) msp_qr_destroy(h_);
  }
  torch::Tensor solve(torch::Tensor B) {
    TORCH_CHECK(B.scalar_type() == torch::kFloat64, "B must be float64");
    TORCH_CHECK(B.dim() == 1 || B.dim() == 2, "B must be a vector or a matrix");
    TORCH_CHECK(B.size(0) == m_, "B row size mismatch: ", B.size(0), " rows, matrix has ", m_);
    const int64_t k = B.dim() == 1 ? 1 : B.size(1);
    const auto dev = torch::Device(torch::kCUDA, device_);
    // column-major right-hand sides on the solver's device
    torch::Tensor Bt = (B.dim() == 1 ? B.reshape({1, m_}) : B.t()).to(dev).contiguous();
    torch::Tensor Xt = torch::empty({k, n_}, torch::TensorOptions().device(dev).dtype(torch::kFloat64));
    {
      c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
      void* s = stream_of(device_);
      py::gil_scoped_release nogil;
      if (k > 0) check(msp_qr_solve(h_, Bt.data_ptr<double>(), m_, k, Xt.data_ptr<double>(), n_, s), "solve");
    }
    torch::Tensor X = B.dim() == 1 ? Xt.reshape({n_}) : Xt.t();
    return X.to(B.device());
  }
  torch::Tensor get_grad_a_helper(torch::Tensor col_starts, torch::Tensor row_indices, torch::Tensor grad_b,
                                  torch::Tensor x, int64_t n_rows, int64_t n_cols) {
    TORCH_CHECK(col_starts.numel() == n_cols + 1, "col_starts: expected ", n_cols + 1, " entries");
    TORCH_CHECK(grad_b.numel() == n_rows && x.numel() == n_cols, "grad_b / x sizes");
    const auto dev = grad_b.device();
    auto cs = col_starts.to(dev, torch::kInt64);
    auto cols = torch::repeat_interleave(torch::arange(n_cols, cs.options()), cs.slice(0, 1) - cs.slice(0, 0, -1));
    auto rows = row_indices.to(dev, torch::kInt64);
    return -(grad_b.to(torch::kFloat64).index_select(0, rows) * x.to(dev, torch::kFloat64).index_select(0, cols));
  }
  py::tuple bandwidth() const {
    int32_t lo = 0, up = 0;
    check(msp_qr_bandwidth(h_, &lo, &up), "bandwidth");
    return py::make_tuple(lo, up);
  }

 private:
  msp_qr h_ = nullptr;
  int64_t m_, n_;
  int device_ = 0;
};

class BatchTensor {
 public:
  BatchTensor(CsrTensor& A, torch::Tensor sys_off) : device_(A.device()) {
    TORCH_CHECK(sys_off.dim() == 1 && sys_off.numel() >= 1, "sys_off must be 1-D with num_systems + 1 entries");
    auto so = sys_off.to(torch::kCPU).to(torch::kInt64).contiguous();
    std::vector<uint32_t> h((size_t)so.numel());
    const int64_t* p = so.data_ptr<int64_t>();
    for (size_t i = 0; i < h.size(); ++i) {
      TORCH_CHECK(p[i] >= 0 && p[i] < (1ll << 32), "sys_off out of range");
      h[i] = (uint32_t)p[i];
    }
    nsys_ = (uint32_t)h.size() - 1;
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    void* s = stream_for(device_);
    py::gil_scoped_release nogil;
    check(msp_batch_create(&h_, A.raw(), h.data(), nsys_, s), "batch_tensor");
  }
  ~BatchTensor() {
    if (h_) msp_batch_destroy(h_);
  }
  int64_t num_systems() const { return nsys_; }
  py::dict pcg(torch::Tensor b, torch::Tensor x, double tol, int64_t maxiter, int64_t steps_per_graph,
               int64_t tile_rows, int64_t rows_per_thread, int64_t max_verify) {
    TORCH_CHECK(b.dim() == 1 || b.dim() == 2, "b must be (n,) or (n, nrhs)");
    const int64_t nrhs = b.dim() == 1 ? 1 : b.size(1);
    TORCH_CHECK(nrhs >= 1 && nrhs <= 4, "1..4 right-hand sides");
    expect(b, torch::kFloat32, -1, "b");
    const bool x64 = x.scalar_type() == torch::kFloat64;
    expect(x, x64 ? torch::kFloat64 : torch::kFloat32, b.numel(), "x");
    msp_batch_opts o;
    msp_default_batch_opts(&o);
    o.tol = tol;
    o.maxiter = (int32_t)maxiter;
    o.nrhs = (int32_t)nrhs;
    o.steps_per_graph = (int32_t)steps_per_graph;
    o.tile_rows = (int32_t)tile_rows;
    o.rows_per_thread = (int32_t)rows_per_thread;
    o.max_verify = (int32_t)max_verify;
    const int64_t np = (int64_t)nsys_ * nrhs;
    auto st = torch::empty({np}, torch::kInt32);
    auto it = torch::empty({np}, torch::kInt32);
    auto rel = torch::empty({np}, torch::kFloat64);
    msp_batch_info info{};
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    void* s = stream_for(device_);
    {
      py::gil_scoped_release nogil;
      if (x64)
        check(msp_batch_pcg_x64(h_, b.data_ptr<float>(), x.data_ptr<double>(), &o, &info, st.data_ptr<int32_t>(),
                                it.data_ptr<int32_t>(), rel.data_ptr<double>(), s),
              "batch pcg");
      else
        check(msp_batch_pcg(h_, b.data_ptr<float>(), x.data_ptr<float>(), &o, &info, st.data_ptr<int32_t>(),
                            it.data_ptr<int32_t>(), rel.data_ptr<double>(), s),
              "batch pcg");
    }
    py::dict d;
    d["converged_pairs"] = info.converged_pairs;
    d["total_pairs"] = info.total_pairs;
    d["max_iterations"] = info.max_iterations;
    d["graph_launches"] = info.graph_launches;
    d["verifications"] = info.verifications;
    d["tile_rows"] = info.tile_rows;
    d["rows_per_thread"] = info.rows_per_thread;
    d["diagonal_planes"] = info.diagonal_planes;
    d["symmetric_storage"] = info.symmetric_storage;
    d["seconds"] = info.seconds;
    d["setup_seconds"] = info.setup_seconds;
    d["status"] = st.view({(int64_t)nsys_, nrhs});
    d["iterations"] = it.view({(int64_t)nsys_, nrhs});
    d["relres"] = rel.view({(int64_t)nsys_, nrhs});
    return d;
  }

 private:
  int device_ = 0;
  uint32_t nsys_ = 0;
  msp_batch h_ = nullptr;
};

// bcsr.bcsr_tensor(keys, values, row_ptr, col_ind, num_rows, num_cols, block_size)
// (iterative/BiCGSTAB/bcsr.cpp:357-422, 953-965): with block arrays in
// values/row_ptr/col_ind (keys unused, as in the reference) or, when row_ptr is
// empty, from packed COO keys converted on the GPU.
class BcsrTensor {
 public:
  BcsrTensor(torch::Tensor keys, torch::Tensor values, torch::Tensor row_ptr, torch::Tensor col_ind,
             int64_t num_rows, int64_t num_cols, int64_t block_size) {
    TORCH_CHECK(num_rows >= 0 && num_cols >= 0 && num_rows < (1ll << 32) && num_cols < (1ll << 32),
                "num_rows / num_cols out of range");
    TORCH_CHECK(block_size == 4 || block_size == 8 || block_size == 16, "block_size must be 4, 8 or 16");
    device_ = pick_device({&keys, &values, &row_ptr, &col_ind});
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    void* s = stream_for(device_);
    const int64_t B = block_size;
    if (row_ptr.defined() && row_ptr.numel() > 0) {
      const int64_t nbr = (num_rows + B - 1) / B;
      expect(row_ptr, torch::kInt32, nbr + 1, "row_ptr (block rows)");
      expect(col_ind, torch::kInt32, -1, "col_ind (block columns)");
      expect(values, torch::kFloat32, col_ind.numel() * B * B, "values (blocks)");
      py::gil_scoped_release nogil;
      check(msp_bcsr_create(&h_, reinterpret_cast<const uint32_t*>(row_ptr.data_ptr<int32_t>()),
                            reinterpret_cast<const uint32_t*>(col_ind.data_ptr<int32_t>()), values.data_ptr<float>(),
                            (uint32_t)num_rows, (uint32_t)num_cols, (uint32_t)B, s),
            "bcsr_tensor");
    } else {
      expect(keys, torch::kInt64, -1, "keys");
      expect(values, torch::kFloat32, keys.numel(), "values");
      py::gil_scoped_release nogil;
      check(msp_bcsr_create_coo(&h_, reinterpret_cast<const uint64_t*>(keys.data_ptr<int64_t>()),
                                values.data_ptr<float>(), (uint64_t)keys.numel(), (uint32_t)num_rows,
                                (uint32_t)num_cols, (uint32_t)B, s),
            "bcsr_tensor (from COO)");
    }
  }
  ~BcsrTensor() {
    if (h_) msp_bcsr_destroy(h_);
  }

  void shape(uint32_t& n, uint32_t& m, uint32_t& B, uint64_t& nb) const {
    check(msp_bcsr_shape(h_, &n, &m, &B, &nb), "shape");
  }

  // vectors may be logical (n / m) or padded to whole blocks
  static void expect_vec(const torch::Tensor& t, int64_t logical, int64_t B, const char* name) {
    const int64_t padded = (logical + B - 1) / B * B;
    expect(t, torch::kFloat32, -1, name);
    TORCH_CHECK(t.numel() == logical || t.numel() == padded, name, ": expected ", logical, " or ", padded,
                " elements, got ", t.numel());
  }

  void mv(torch::Tensor x, torch::Tensor b) {
    uint32_t n, m, B;
    uint64_t nb;
    shape(n, m, B, nb);
    expect_vec(x, m, B, "x");
    expect_vec(b, n, B, "b");
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    void* s = stream_for(device_);
    py::gil_scoped_release nogil;
    check(msp_bcsr_spmv(h_, x.data_ptr<float>(), b.data_ptr<float>(), s), "mv");
  }

  py::dict solve(bool bicg, torch::Tensor b, torch::Tensor x, double tol, int64_t maxiter, int64_t check_every,
                 const std::string& criterion, bool use_x0, const std::string& precond, bool verify) {
    uint32_t n, m, B;
    uint64_t nb;
    shape(n, m, B, nb);
    expect_vec(b, n, B, "b");
    expect_vec(x, m, B, "x");
    msp_solve_opts o;
    msp_default_opts(&o);
    o.tol = tol;
    o.maxiter = (int32_t)maxiter;
    o.check_every = (int32_t)check_every;
    TORCH_CHECK(criterion == "relres" || criterion == "reference", "criterion must be 'relres' or 'reference'");
    o.criterion = criterion == "relres" ? MSP_CRIT_RELRES : MSP_CRIT_REFERENCE;
    o.use_x0 = use_x0 ? 1 : 0;
    TORCH_CHECK(precond == "none" || precond == "jacobi", "precond must be 'none' or 'jacobi'");
    o.precond = precond == "jacobi" ? MSP_PRECOND_JACOBI : MSP_PRECOND_NONE;
    o.verify = verify ? 1 : 0;
    msp_solve_info info{};
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    void* s = stream_for(device_);
    {
      py::gil_scoped_release nogil;
      if (bicg) check(msp_bcsr_bicgstab(h_, b.data_ptr<float>(), x.data_ptr<float>(), &o, &info, s), "bicgstab");
      else check(msp_bcsr_cg(h_, b.data_ptr<float>(), x.data_ptr<float>(), &o, &info, s), "cgd");
    }
    return info_dict(info);
  }

  void richardson(torch::Tensor b, torch::Tensor x, double omega, int64_t iters, bool use_x0) {
    uint32_t n, m, B;
    uint64_t nb;
    shape(n, m, B, nb);
    expect_vec(b, n, B, "b");
    expect_vec(x, m, B, "x");
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    void* s = stream_for(device_);
    py::gil_scoped_release nogil;
    check(msp_bcsr_richardson(h_, b.data_ptr<float>(), x.data_ptr<float>(), (float)omega, (int32_t)iters,
                              use_x0 ? 1 : 0, s),
          "richardson");
  }

  py::tuple export_arrays() {
    uint32_t n, m, B;
    uint64_t nb;
    shape(n, m, B, nb);
    const int64_t nbr = (n + B - 1) / B;
    auto opt = torch::TensorOptions().device(torch::kCUDA, device_);
    torch::Tensor rp = torch::empty({nbr + 1}, opt.dtype(torch::kInt32));
    torch::Tensor bc = torch::empty({(int64_t)nb}, opt.dtype(torch::kInt32));
    torch::Tensor bl = torch::empty({(int64_t)nb, (int64_t)B, (int64_t)B}, opt.dtype(torch::kFloat32));
    c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
    void* s = stream_for(device_);
    {
      py::gil_scoped_release nogil;
      check(msp_bcsr_export(h_, reinterpret_cast<uint32_t*>(rp.data_ptr<int32_t>()),
                            nb ? reinterpret_cast<uint32_t*>(bc.data_ptr<int32_t>()) : nullptr,
                            nb ? bl.data_ptr<float>() : nullptr, s),
            "export");
    }
    return py::make_tuple(rp, bc, bl);
  }

  int64_t num_rows() const { uint32_t n, m, B; uint64_t nb; shape(n, m, B, nb); return n; }
  int64_t num_cols() const { uint32_t n, m, B; uint64_t nb; shape(n, m, B, nb); return m; }
  int64_t block_size() const { uint32_t n, m, B; uint64_t nb; shape(n, m, B, nb); return B; }
  int64_t num_blocks() const { uint32_t n, m, B; uint64_t nb; shape(n, m, B, nb); return (int64_t)nb; }
  uintptr_t handle() const { return reinterpret_cast<uintptr_t>(h_); }

 private:
  msp_bcsr h_ = nullptr;
  int device_ = 0;
};

void coo_csr(torch::Tensor keys, torch::Tensor values, torch::Tensor row_ptr, torch::Tensor col_ind,
             torch::Tensor out_vals, int64_t num_rows, int64_t num_cols) {
  TORCH_CHECK(num_rows >= 0 && num_cols >= 0 && num_rows < (1ll << 32) && num_cols < (1ll << 32),
              "num_rows / num_cols out of range");
  expect(keys, torch::kInt64, -1, "keys");
  expect(values, torch::kFloat32, keys.numel(), "values");
  expect(row_ptr, torch::kInt32, num_rows + 1, "row_ptr");
  expect(col_ind, torch::kInt32, keys.numel(), "col_ind");
  expect(out_vals, torch::kFloat32, keys.numel(), "out_vals");
  int dev = pick_device({&keys, &values, &row_ptr, &col_ind, &out_vals});
  c10::hip::HIPGuard guard((c10::DeviceIndex)dev);
  void* s = stream_for(dev);
  py::gil_scoped_release nogil;
  check(msp_coo_to_csr(reinterpret_cast<const uint64_t*>(keys.data_ptr<int64_t>()), values.data_ptr<float>(),
                       (uint64_t)keys.numel(), (uint32_t)num_rows, (uint32_t)num_cols,
                       reinterpret_cast<uint32_t*>(row_ptr.data_ptr<int32_t>()),
                       reinterpret_cast<uint32_t*>(col_ind.data_ptr<int32_t>()), out_vals.data_ptr<float>(), s),
        "coo_csr");
}

}  // namespace

void mlp_forward(torch::Tensor params, torch::Tensor feat, torch::Tensor out) {
  expect(params, torch::kFloat32, msp_mlp_num_params(), "params");
  expect(feat, torch::kFloat32, -1, "feat");
  TORCH_CHECK(feat.dim() == 2 && feat.size(1) == 8, "feat must be (num_edges, 8)");
  expect(out, torch::kFloat32, feat.size(0), "out");
  c10::hip::HIPGuard guard(params.device().index());
  check(msp_mlp_forward(params.data_ptr<float>(), feat.data_ptr<float>(), (uint64_t)feat.size(0), out.data_ptr<float>(),
                        stream_of(params.device().index())),
        "mlp_forward");
}

void mlp_backward(torch::Tensor params, torch::Tensor feat, torch::Tensor go, torch::Tensor grad) {
  expect(params, torch::kFloat32, msp_mlp_num_params(), "params");
  expect(feat, torch::kFloat32, -1, "feat");
  TORCH_CHECK(feat.dim() == 2 && feat.size(1) == 8, "feat must be (num_edges, 8)");
  expect(go, torch::kFloat32, feat.size(0), "grad_out");
  expect(grad, torch::kFloat32, msp_mlp_num_params(), "grad_params");
  c10::hip::HIPGuard guard(params.device().index());
  check(msp_mlp_backward(params.data_ptr<float>(), feat.data_ptr<float>(), go.data_ptr<float>(), (uint64_t)feat.size(0),
                         grad.data_ptr<float>(), stream_of(params.device().index())),
        "mlp_backward");
}

void mlp_adam(torch::Tensor params, torch::Tensor grad, torch::Tensor m, torch::Tensor v, double lr, double b1,
              double b2, double eps, int64_t step, double wd) {
  for (auto* t : {&params, &grad, &m, &v}) expect(*t, torch::kFloat32, msp_mlp_num_params(), "adam tensor");
  c10::hip::HIPGuard guard(params.device().index());
  check(msp_mlp_adam(params.data_ptr<float>(), grad.data_ptr<float>(), m.data_ptr<float>(), v.data_ptr<float>(),
                     (float)lr, (float)b1, (float)b2, (float)eps, (int32_t)step, (float)wd,
                     stream_of(params.device().index())),
        "mlp_adam");
}

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.attr("abi_version") = msp_abi_version();
  m.attr("mlp_num_params") = msp_mlp_num_params();
  m.def("mlp_forward", &mlp_forward, "edge MLP forward (BF16 MFMA)", py::arg("params"), py::arg("feat"),
        py::arg("out"));
  m.def("mlp_backward", &mlp_backward, "edge MLP backward: grad = d(sum grad_out*out)/d params", py::arg("params"),
        py::arg("feat"), py::arg("grad_out"), py::arg("grad"));
  m.def("mlp_adam", &mlp_adam, "Adam step", py::arg("params"), py::arg("grad"), py::arg("m"), py::arg("v"),
        py::arg("lr"), py::arg("beta1") = 0.9, py::arg("beta2") = 0.999, py::arg("eps") = 1e-8, py::arg("step") = 1,
        py::arg("weight_decay") = 0.0);
  m.def("pcg_fsai", &pcg_fsai, "FSAI-preconditioned CG", py::arg("A"), py::arg("F"), py::arg("b"), py::arg("x"),
        py::arg("tol") = 1e-6, py::arg("maxiter") = 10000, py::arg("check_every") = 16, py::arg("use_x0") = false,
        py::arg("values") = "f32", py::arg("verify") = true);
  m.def("pcg_fsai_refined", &pcg_fsai_refined,
        "FP64 iterative refinement around FSAI-PCG corrections (b, x float64)", py::arg("A"), py::arg("F"),
        py::arg("b"), py::arg("x"), py::arg("tol") = 1e-6, py::arg("inner_tol") = 0.0, py::arg("max_outer") = 30,
        py::arg("inner_maxiter") = 10000, py::arg("values") = "f32", py::arg("use_x0") = false);
  m.def("device_count", []() { return msp_device_count(); });
  m.def(
      "release_workspace", []() { check(msp_release_workspace(), "release_workspace"); },
      "return the cached temporaries of matrix builds (a scratch arena per device and stream) to the system");
  m.def("coo_csr", &coo_csr, "COO -> CSR (stable packed-key sort)", py::arg("keys"), py::arg("values"),
        py::arg("row_ptr"), py::arg("col_ind"), py::arg("out_vals"), py::arg("num_rows"), py::arg("num_cols"));
  py::class_<CsrTensor>(m, "csr_tensor")
      .def(py::init<torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, int64_t, int64_t>(),
           "CSR Tensor Constructor", py::arg("keys"), py::arg("values"), py::arg("row_ptr"), py::arg("col_ind"),
           py::arg("out_vals"), py::arg("num_rows"), py::arg("num_cols"))
      .def("mv", &CsrTensor::mv, "b = A x", py::arg("x"), py::arg("b"), py::arg("values") = "f32")
      .def(
          "cgd",
          [](CsrTensor& self, torch::Tensor b, torch::Tensor x, double tol, int64_t maxiter, int64_t check_every,
             const std::string& criterion, bool use_x0, const std::string& precond, const std::string& values,
             bool verify, const std::string& cg_form) {
            return self.solve(false, b, x, tol, maxiter, check_every, criterion, use_x0, precond, values, verify,
                              cg_form);
          },
          "conjugate gradient", py::arg("b"), py::arg("x"), py::arg("tol") = 1e-6, py::arg("maxiter") = 10000,
          py::arg("check_every") = 16, py::arg("criterion") = "relres", py::arg("use_x0") = false,
          py::arg("precond") = "none", py::arg("values") = "f32", py::arg("verify") = true,
          py::arg("cg_form") = "auto")
      .def(
          "bicgstab",
          [](CsrTensor& self, torch::Tensor b, torch::Tensor x, double tol, int64_t maxiter, int64_t check_every,
             const std::string& criterion, bool use_x0, const std::string& values, bool verify) {
            return self.solve(true, b, x, tol, maxiter, check_every, criterion, use_x0, "none", values, verify);
          },
          "BiCGSTAB", py::arg("b"), py::arg("x"), py::arg("tol") = 1e-6, py::arg("maxiter") = 10000,
          py::arg("check_every") = 16, py::arg("criterion") = "relres", py::arg("use_x0") = false,
          py::arg("values") = "f32", py::arg("verify") = true)
      .def("richardson", &CsrTensor::richardson, "x <- x + omega (b - A x), iters times", py::arg("b"), py::arg("x"),
           py::arg("omega") = 0.01, py::arg("iters") = 1000, py::arg("use_x0") = false, py::arg("values") = "f32")
      .def("solve_refined", &CsrTensor::solve_refined,
           "FP64 iterative refinement around FP32 (BF16/FP16-valued) Krylov corrections; b, x float64",
           py::arg("b"), py::arg("x"), py::arg("tol") = 1e-10, py::arg("inner_tol") = 0.0,
           py::arg("max_outer") = 30, py::arg("inner_maxiter") = 10000, py::arg("solver") = "cg",
           py::arg("values") = "f32", py::arg("precond") = "none", py::arg("use_x0") = false)
      .def("export", &CsrTensor::export_into, "write row_ptr / col_ind / values", py::arg("row_ptr"),
           py::arg("col_ind"), py::arg("out_vals"))
      .def("set_format", &CsrTensor::set_format,
           "'auto' (diagonal storage for stencil-like patterns, the default) or 'csr'", py::arg("format"))
      .def_property_readonly("format", &CsrTensor::format, "storage the SpMV uses: 'dia' or 'csr'")
      .def_property_readonly("num_rows", &CsrTensor::num_rows)
      .def_property_readonly("num_cols", &CsrTensor::num_cols)
      .def_property_readonly("nnz", &CsrTensor::nnz)
      .def_property_readonly("device", &CsrTensor::device)
      .def_property_readonly("handle", &CsrTensor::handle);
  py::class_<BatchTensor>(m, "batch_tensor")
      .def(py::init<CsrTensor&, torch::Tensor>(), py::arg("A"), py::arg("sys_off"), py::keep_alive<1, 2>())
      .def_property_readonly("num_systems", &BatchTensor::num_systems)
      .def("pcg", &BatchTensor::pcg,
           "Jacobi-PCG of every (system, rhs) pair (x float32, or float64 for the compensated iterate's exact "
           "sum); status codes 1 converged, 2 breakdown, 3 maxiter, 4 stagnated",
           py::arg("b"), py::arg("x"), py::arg("tol") = 1e-6, py::arg("maxiter") = 10000,
           py::arg("steps_per_graph") = 16, py::arg("tile_rows") = 0, py::arg("rows_per_thread") = 0,
           py::arg("max_verify") = 3);
  py::class_<QrSolver>(m, "SparseQRSolver")
      .def(py::init<torch::Tensor, torch::Tensor, torch::Tensor, int64_t, int64_t>(), py::arg("col_starts"),
           py::arg("row_indices"), py::arg("values"), py::arg("n_rows"), py::arg("n_cols"))
      .def("solve", &QrSolver::solve, py::arg("B"))
      .def("get_grad_a_helper", &QrSolver::get_grad_a_helper, py::arg("col_starts"), py::arg("row_indices"),
           py::arg("grad_b"), py::arg("x"), py::arg("n_rows"), py::arg("n_cols"))
      .def_property_readonly("bandwidth", &QrSolver::bandwidth);
  py::class_<FsaiTensor>(m, "fsai_tensor")
      .def(py::init<CsrTensor&>(), py::arg("A"), py::keep_alive<1, 2>())
      .def_property_readonly("num_edges", &FsaiTensor::num_edges)
      .def("edges", &FsaiTensor::edges)
      .def("features", &FsaiTensor::features)
      .def("classical", &FsaiTensor::classical, "classical FSAI values; returns the number of Jacobi fallback rows")
      .def("set_values", &FsaiTensor::set_values, py::arg("g"))
      .def("select_features", &FsaiTensor::select_features, py::arg("feature_set"),
           "edge-network inputs: 'coupled8' (default) or 'stated5' (a_ij, a_ii, a_jj, row nnz, row |sum|; "
           "G's rows normalised to diag(G A G^T) = 1)")
      .def_property_readonly("feature_set", &FsaiTensor::feature_set)
      .def("set_network", &FsaiTensor::set_network, py::arg("params"))
      .def("set_network_output", &FsaiTensor::set_network_output, py::arg("out"))
      .def("values", &FsaiTensor::values)
      .def("apply", &FsaiTensor::apply, py::arg("r"), py::arg("z"))
      .def("probe_loss_grad", &FsaiTensor::probe_loss_grad, py::arg("W"), py::arg("grad"))
      .def("train_step", &FsaiTensor::train_step, py::arg("params"), py::arg("W"), py::arg("grad_params"));
  py::class_<BcsrTensor>(m, "bcsr_tensor")
      .def(py::init<torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, int64_t, int64_t, int64_t>(),
           "BCSR Tensor Constructor", py::arg("keys"), py::arg("values"), py::arg("row_ptr"), py::arg("col_ind"),
           py::arg("num_rows"), py::arg("num_cols"), py::arg("block_size") = 4)
      .def("mv", &BcsrTensor::mv, "b = A x", py::arg("x"), py::arg("b"))
      .def(
          "cgd",
          [](BcsrTensor& self, torch::Tensor b, torch::Tensor x, double tol, int64_t maxiter, int64_t check_every,
             const std::string& criterion, bool use_x0, const std::string& precond, bool verify) {
            return self.solve(false, b, x, tol, maxiter, check_every, criterion, use_x0, precond, verify);
          },
          "conjugate gradient", py::arg("b"), py::arg("x"), py::arg("tol") = 1e-6, py::arg("maxiter") = 10000,
          py::arg("check_every") = 16, py::arg("criterion") = "relres", py::arg("use_x0") = false,
          py::arg("precond") = "none", py::arg("verify") = true)
      .def(
          "bicgstab",
          [](BcsrTensor& self, torch::Tensor b, torch::Tensor x, double tol, int64_t maxiter, int64_t check_every,
             const std::string& criterion, bool use_x0, bool verify) {
            return self.solve(true, b, x, tol, maxiter, check_every, criterion, use_x0, "none", verify);
          },
          "BiCGSTAB", py::arg("b"), py::arg("x"), py::arg("tol") = 1e-6, py::arg("maxiter") = 10000,
          py::arg("check_every") = 16, py::arg("criterion") = "relres", py::arg("use_x0") = false,
          py::arg("verify") = true)
      .def("richardson", &BcsrTensor::richardson, "x <- x + omega (b - A x), iters times", py::arg("b"),
           py::arg("x"), py::arg("omega") = 0.01, py::arg("iters") = 1000, py::arg("use_x0") = false)
      .def("export", &BcsrTensor::export_arrays, "(block row_ptr, block col_ind, blocks) as device tensors")
      .def_property_readonly("num_rows", &BcsrTensor::num_rows)
      .def_property_readonly("num_cols", &BcsrTensor::num_cols)
      .def_property_readonly("block_size", &BcsrTensor::block_size)
      .def_property_readonly("num_blocks", &BcsrTensor::num_blocks)
      .def_property_readonly("handle", &BcsrTensor::handle);
}
