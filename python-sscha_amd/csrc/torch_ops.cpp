// PyTorch-ROCm operators over libtdscha_hip (the boundary the Python host calls on the hot
// path): torch.ops.tdscha.* take device tensors, check their dtype, device, layout and size
// before any launch, and enqueue the library's kernels on the caller's current HIP stream.
// Ensembles and projectors are passed as the integer values of their C-ABI handles
// (include/tdscha_hip.h), created and owned by the Python objects that wrap them.
#include <ATen/hip/HIPContext.h>
#include <c10/hip/HIPGuard.h>
#include <torch/library.h>

#include <string>
#include <vector>

#include "../../include/tdscha_hip.h"

namespace {

void check_rc(int rc, const char *what) {
  if (rc == 0) return;
  const std::string msg = std::string(what) + ": " + tdscha_last_error();
  if (rc >= 1 && rc <= 5) TORCH_CHECK_VALUE(false, msg);
  TORCH_CHECK(false, msg, " (code ", rc, ")");
}

void check_vec(const at::Tensor &t, int64_t numel, const at::Device &dev, const char *what) {
  TORCH_CHECK_VALUE(t.is_cuda() && t.device() == dev, what, " must be on ", dev, ", got ",
                    t.device());
  TORCH_CHECK_VALUE(t.scalar_type() == at::kDouble, what, " must be float64, got ",
                    t.scalar_type());
  TORCH_CHECK_VALUE(t.is_contiguous(), what, " must be contiguous");
  TORCH_CHECK_VALUE(t.numel() == numel, what, " must have ", numel, " entries, got ", t.numel());
}

struct Dims {
  int n, N, N_sym, s_batch, npd, Np;
};

Dims dims(int64_t ens) {
  int d[6];
  check_rc(tdscha_ensemble_dims(reinterpret_cast<tdscha_ensemble *>(ens), d),
           "tdscha_ensemble_dims");
  return Dims{d[0], d[1], d[2], d[3], d[4], d[5]};
}

void *stream_of(const at::Device &dev) {
  return reinterpret_cast<void *>(c10::hip::getCurrentHIPStream(dev.index()).stream());
}

void apply_full_L_batched(int64_t ens, int64_t precision, at::TensorList psis,
                          at::IntArrayRef flags, at::ArrayRef<double> shifts,
                          at::TensorList outs) {
  const int C = (int)psis.size();
  TORCH_CHECK_VALUE(C >= 1 && C <= TDSCHA_MAX_CHAINS, "between 1 and ", TDSCHA_MAX_CHAINS,
                    " vectors per call");
  TORCH_CHECK_VALUE((int)outs.size() == C && (int)flags.size() == C && (int)shifts.size() == C,
                    "psis, flags, shifts and outs must have the same length");
  const Dims d = dims(ens);
  const int64_t n_psi = (int64_t)d.n + (int64_t)d.n * (d.n + 1);
  const at::Device dev = psis[0].device();
  std::vector<const double *> in(C);
  std::vector<double *> out(C);
  std::vector<int> fl(C);
  for (int c = 0; c < C; ++c) {
    check_vec(psis[c], n_psi, dev, "psi");
    check_vec(outs[c], n_psi, dev, "out");
    in[c] = psis[c].data_ptr<double>();
    out[c] = outs[c].data_ptr<double>();
    fl[c] = (int)flags[c];
  }
  for (int c = 0; c < C; ++c)
    for (int k = 0; k < C; ++k)
      TORCH_CHECK_VALUE(out[c] != in[k], "an output must not alias an input");
  const c10::hip::HIPGuard guard(dev.index());
  check_rc(tdscha_apply_full_L_batched(reinterpret_cast<tdscha_ensemble *>(ens), (int)precision,
                                       C, fl.data(), shifts.data(), in.data(), out.data(),
                                       stream_of(dev)),
           "tdscha_apply_full_L_batched");
}

void ensemble_apply(int64_t ens, int64_t precision, int64_t terms, const at::Tensor &R1,
                    const at::Tensor &Y1, const at::Tensor &f, const at::Tensor &d2v,
                    bool accumulate) {
  const Dims d = dims(ens);
  const at::Device dev = R1.device();
  check_vec(R1, d.n, dev, "R1");
  check_vec(Y1, (int64_t)d.n * d.n, dev, "Y1");
  check_vec(f, d.n, dev, "f");
  check_vec(d2v, (int64_t)d.n * d.n, dev, "d2v");
  const c10::hip::HIPGuard guard(dev.index());
  check_rc(tdscha_ensemble_apply(reinterpret_cast<tdscha_ensemble *>(ens), (int)precision,
                                 (int)terms, R1.data_ptr<double>(), Y1.data_ptr<double>(),
                                 f.data_ptr<double>(), d2v.data_ptr<double>(), accumulate ? 1 : 0,
                                 stream_of(dev)),
           "tdscha_ensemble_apply");
}

void force_grad(int64_t ens, int64_t precision, at::TensorList us, at::TensorList vs,
                at::IntArrayRef flags, const at::Tensor &grad, bool accumulate) {
  const int C = (int)us.size();
  TORCH_CHECK_VALUE(C >= 1 && C <= TDSCHA_MAX_CHAINS, "between 1 and ", TDSCHA_MAX_CHAINS,
                    " pairs per call");
  TORCH_CHECK_VALUE((int)vs.size() == C && (int)flags.size() == C,
                    "us, vs and flags must have the same length");
  const Dims d = dims(ens);
  const int64_t n_psi = (int64_t)d.n + (int64_t)d.n * (d.n + 1);
  const at::Device dev = grad.device();
  check_vec(grad, (int64_t)d.N * d.n, dev, "grad");
  std::vector<const double *> pu(C), pv(C);
  std::vector<int> fl(C);
  for (int c = 0; c < C; ++c) {
    check_vec(us[c], n_psi, dev, "u");
    check_vec(vs[c], n_psi, dev, "v");
    pu[c] = us[c].data_ptr<double>();
    pv[c] = vs[c].data_ptr<double>();
    fl[c] = (int)flags[c];
  }
  const c10::hip::HIPGuard guard(dev.index());
  check_rc(tdscha_ensemble_force_grad(reinterpret_cast<tdscha_ensemble *>(ens), (int)precision, C,
                                      fl.data(), pu.data(), pv.data(), grad.data_ptr<double>(),
                                      accumulate ? 1 : 0, stream_of(dev)),
           "tdscha_ensemble_force_grad");
}

void projector_apply(int64_t proj, int64_t n_modes, bool adjoint, const at::Tensor &psi,
                     const at::Tensor &out) {
  const int64_t n_psi = n_modes + n_modes * (n_modes + 1);
  const at::Device dev = psi.device();
  check_vec(psi, n_psi, dev, "psi");
  check_vec(out, n_psi, dev, "out");
  const c10::hip::HIPGuard guard(dev.index());
  check_rc(tdscha_projector_apply(reinterpret_cast<tdscha_projector *>(proj), adjoint ? 1 : 0,
                                  psi.data_ptr<double>(), out.data_ptr<double>(), stream_of(dev)),
           "tdscha_projector_apply");
}

}  // namespace

TORCH_LIBRARY(tdscha, m) {
  m.def("apply_full_L_batched(int ens, int precision, Tensor[] psis, int[] flags, "
        "float[] shifts, Tensor[] outs) -> ()");
  m.def("ensemble_apply(int ens, int precision, int terms, Tensor R1, Tensor Y1, Tensor f, "
        "Tensor d2v, bool accumulate) -> ()");
  m.def("force_grad(int ens, int precision, Tensor[] us, Tensor[] vs, int[] flags, "
        "Tensor grad, bool accumulate) -> ()");
  m.def("projector_apply(int proj, int n_modes, bool adjoint, Tensor psi, Tensor out) -> ()");
}

TORCH_LIBRARY_IMPL(tdscha, CompositeExplicitAutograd, m) {
  m.impl("apply_full_L_batched", &apply_full_L_batched);
  m.impl("ensemble_apply", &ensemble_apply);
  m.impl("force_grad", &force_grad);
  m.impl("projector_apply", &projector_apply);
}
