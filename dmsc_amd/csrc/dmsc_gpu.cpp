// Thin PyTorch-ROCm binding of the C-ABI (include/dmsc_hip.h): module dmsc_amd.csrc.dmsc_gpu.
//
// Mirrors the reference binding dmsc.csrc.dmsc_gpu (dmsc/csrc/dmsc_gpu.cpp:128-254): same function
// name, argument names, defaults and error texts, and a DMSComplex object with 19 read-only fields,
// __getitem__ and __len__. Outputs are torch tensors created by the allocator callback on the
// input's device (shape is a CPU int32[2], as in the reference GPU path); kernels run on the
// current HIP stream of that device.
#include <c10/hip/HIPGuard.h>
#include <c10/hip/HIPStream.h>
#include <pybind11/stl.h>
#include <torch/extension.h>

#include <limits>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "dmsc_hip.h"

namespace {

struct DMSComplex {
  torch::Tensor t[DMSC_NUM_OUTPUTS];
};

const char* kNames[DMSC_NUM_OUTPUTS] = {
    "shape",   "max_pts",    "min_pts",    "sad_pts",     "e_max",          "e_min",
    "p_max",   "p_min",      "ppairs_max", "ppairs_min",  "grad_indices",   "peaks",
    "ridges",  "ridge_offsets", "ridge_arc_offsets", "basins", "valleys", "valley_offsets",
    "valley_arc_offsets"};

struct AllocUser {
  torch::Device dev;
  c10::hip::HIPStream stream;         // the caller's stream: outputs belong to it
  DMSComplex* res;                    // one per field
  std::vector<torch::Tensor> shared;  // buffers holding one slot of several fields (field = -1)
};

torch::TensorOptions opts_for(int32_t dtype, torch::Device dev) {
  return torch::TensorOptions().dtype(dtype == DMSC_DTYPE_FLOAT32 ? torch::kFloat32 : torch::kInt32).device(dev);
}

// The library calls this one call at a time, for a large batch partly from its worker threads: the
// guard makes every output a block of the caller's stream in the caching allocator, whichever thread
// asks for it.
void* alloc_cb(void* user, int32_t field, int32_t slot, int32_t ndim, const int64_t* dims, int32_t dtype) {
  auto* a = static_cast<AllocUser*>(user);
  c10::hip::HIPStreamGuard guard(a->stream);
  std::vector<int64_t> shape(dims, dims + ndim);
  if (field < 0) {
    a->shared.push_back(torch::empty(shape, opts_for(dtype, a->dev)));
    return a->shared.back().data_ptr();
  }
  torch::Tensor& t = a->res[field].t[slot];
  t = torch::empty(shape, opts_for(dtype, a->dev));
  return t.data_ptr();
}

// A contiguous view of `shape` at element offset `off` of `buf`, made directly on buf's storage
// (sharing its version counter). A batch makes thousands of these; the dispatcher's narrow + view
// pair costs a few microseconds each, this a fraction of one.
torch::Tensor view_of(const torch::Tensor& buf, int64_t off, c10::IntArrayRef shape) {
  auto impl = c10::make_intrusive<c10::TensorImpl>(c10::Storage(buf.storage()), buf.key_set(), buf.dtype());
  impl->set_sizes_contiguous(shape);
  impl->set_storage_offset(buf.storage_offset() + off);
  impl->set_version_counter(buf.unsafeGetTensorImpl()->version_counter());
  return torch::Tensor(std::move(impl));
}

// Outputs that live in a shared buffer become views of it (each field's slice, in its own shape).
void attach_views(AllocUser& u, const dmsc_outputs& out, DMSComplex& r) {
  for (int k = 1; k < DMSC_NUM_OUTPUTS; ++k) {
    if (r.t[k].defined()) continue;
    const c10::IntArrayRef shape(out.dims[k], out.ndim[k]);
    int64_t numel = 1;
    for (int64_t d : shape) numel *= d;
    if (numel == 0 || !out.data[k]) {
      r.t[k] = torch::empty(shape, opts_for(out.dtype[k], u.dev));
      continue;
    }
    const char* p = static_cast<const char*>(out.data[k]);
    bool found = false;
    for (const torch::Tensor& buf : u.shared) {
      const char* b = static_cast<const char*>(buf.data_ptr());
      const int64_t bytes = buf.numel() * buf.element_size();
      if (b && p >= b && p + numel * buf.element_size() <= b + bytes) {
        TORCH_CHECK((p - b) % buf.element_size() == 0 && buf.scalar_type() == opts_for(out.dtype[k], u.dev).dtype(),
                    "dmsc: output slot ", k, " does not match its buffer's type or alignment");
        r.t[k] = view_of(buf, (p - b) / buf.element_size(), shape);
        found = true;
        break;
      }
    }
    TORCH_CHECK(found, "dmsc: output slot ", k, " does not point into an allocated buffer");
  }
}

torch::Tensor shape_tensor(const dmsc_outputs& out) {
  torch::Tensor s = torch::empty({2}, torch::TensorOptions().dtype(torch::kInt32));
  s.data_ptr<int32_t>()[0] = out.shape[0];
  s.data_ptr<int32_t>()[1] = out.shape[1];
  return s;
}

dmsc_context* context_for(int device) {
  static std::mutex mu;
  static std::unordered_map<int, dmsc_context*> ctxs;  // one per device, kept for the process
  std::lock_guard<std::mutex> lock(mu);
  auto it = ctxs.find(device);
  if (it != ctxs.end()) return it->second;
  dmsc_context* c = nullptr;
  TORCH_CHECK(dmsc_context_create(device, &c) == DMSC_OK, "dmsc: could not create a HIP context on device ", device);
  ctxs[device] = c;
  return c;
}

// reference: dmsc/csrc/input_validation.hxx:8-27
void validate(const torch::Tensor& f) {
  TORCH_CHECK(f.defined(), "Input tensor must be defined");
  TORCH_CHECK(f.layout() == c10::kStrided, "Input tensor must use strided layout");
  TORCH_CHECK(f.scalar_type() == torch::kFloat32, "Input tensor must have dtype torch.float32, got ", f.scalar_type());
  TORCH_CHECK(f.dim() == 2 || f.dim() == 3, "Input tensor must be 2D [H, W] or 3D [B, H, W], got ", f.dim(), "D");
  if (f.dim() == 3) TORCH_CHECK(f.size(0) > 0, "Batch dimension must be non-empty");
  const int64_t h = f.size(-2), w = f.size(-1);
  TORCH_CHECK(h > 0 && w > 0, "Spatial dimensions must be non-empty, got ", h, "x", w);
  const int64_t max_cells = std::numeric_limits<int32_t>::max();
  TORCH_CHECK(h + 1 <= max_cells / 4 / (w + 1), "Input is too large for 32-bit cell IDs: ", h, "x", w);
}

uint32_t make_flags(bool grad, bool dual, bool tmax, bool tmin, bool gmax, bool gmin) {
  return (grad ? DMSC_RETURN_GRADIENT : 0u) | (dual ? DMSC_DUAL : 0u) | (tmax ? DMSC_TRACE_MAX_ARCS : 0u) |
         (tmin ? DMSC_TRACE_MIN_ARCS : 0u) | (gmax ? DMSC_TRACE_MAX_GROUPS : 0u) | (gmin ? DMSC_TRACE_MIN_GROUPS : 0u);
}

DMSComplex run_field(dmsc_context* ctx, const torch::Tensor& f, float thr, uint32_t flags,
                     const c10::hip::HIPStream& stream) {
  DMSComplex r;
  AllocUser u{f.device(), stream, &r, {}};
  dmsc_outputs out;
  dmsc_status st;
  {
    pybind11::gil_scoped_release release;
    st = dmsc_extract(ctx, f.data_ptr<float>(), (int32_t)f.size(0), (int32_t)f.size(1), thr, flags,
                      (dmsc_stream_t)stream.stream(), alloc_cb, &u, &out);
  }
  TORCH_CHECK(st == DMSC_OK, dmsc_last_error(ctx));
  r.t[0] = shape_tensor(out);
  return r;
}

// Fields of one device processed together by the C-ABI batch entry (one pipeline for all fields).
std::vector<DMSComplex> run_batch(dmsc_context* ctx, torch::Device dev, const std::vector<const float*>& ptrs,
                                  const std::vector<int32_t>& hs, const std::vector<int32_t>& ws,
                                  const std::vector<float>& thr, uint32_t flags, const c10::hip::HIPStream& stream) {
  const int B = (int)ptrs.size();
  std::vector<DMSComplex> res(B);
  AllocUser u{dev, stream, res.data(), {}};
  std::vector<dmsc_outputs> outs(B);
  dmsc_status st;
  {
    pybind11::gil_scoped_release release;
    st = dmsc_extract_batch(ctx, B, ptrs.data(), hs.data(), ws.data(), thr.data(), flags,
                            (dmsc_stream_t)stream.stream(), alloc_cb, &u, outs.data());
  }
  TORCH_CHECK(st == DMSC_OK, dmsc_last_error(ctx));
  torch::Tensor shapes = torch::empty({B, 2}, torch::TensorOptions().dtype(torch::kInt32));
  int32_t* sp = shapes.data_ptr<int32_t>();
  for (int b = 0; b < B; ++b) {
    attach_views(u, outs[b], res[b]);
    sp[2 * b] = outs[b].shape[0];
    sp[2 * b + 1] = outs[b].shape[1];
    res[b].t[0] = view_of(shapes, 2 * b, {2});
  }
  return res;
}

pybind11::object extract_dmsc(torch::Tensor scalar_field, float persistence_threshold, bool return_gradient,
                              bool is_dual, bool trace_max_arcs, bool trace_min_arcs, bool trace_max_groups,
                              bool trace_min_groups) {
  validate(scalar_field);
  TORCH_CHECK(scalar_field.device().is_cuda(), "GPU backend requires a CUDA or MPS tensor, got ",
              scalar_field.device());
  c10::hip::HIPGuard guard(scalar_field.device().index());
  scalar_field = scalar_field.contiguous();
  dmsc_context* ctx = context_for(scalar_field.device().index());
  const c10::hip::HIPStream stream = c10::hip::getCurrentHIPStream(scalar_field.device().index());
  const uint32_t flags = make_flags(return_gradient, is_dual, trace_max_arcs, trace_min_arcs, trace_max_groups,
                                    trace_min_groups);
  if (scalar_field.dim() == 2) return pybind11::cast(run_field(ctx, scalar_field, persistence_threshold, flags, stream));
  const int64_t B = scalar_field.size(0), H = scalar_field.size(1), W = scalar_field.size(2);
  TORCH_CHECK(B <= std::numeric_limits<int32_t>::max(), "batch too large");
  std::vector<const float*> ptrs(B);
  for (int64_t b = 0; b < B; ++b) ptrs[b] = scalar_field.data_ptr<float>() + b * H * W;
  std::vector<int32_t> hs(B, (int32_t)H), ws(B, (int32_t)W);
  std::vector<float> thr(B, persistence_threshold);
  return pybind11::cast(run_batch(ctx, scalar_field.device(), ptrs, hs, ws, thr, flags, stream));
}

// Mixed-size batch: a list of 2-D fields with one threshold each (no reference counterpart; the
// reference batches only one shape with one threshold, dmsc_gpu.cpp:84-126).
std::vector<DMSComplex> extract_dmsc_list(std::vector<torch::Tensor> fields, std::vector<double> thresholds,
                                          bool return_gradient, bool is_dual, bool trace_max_arcs,
                                          bool trace_min_arcs, bool trace_max_groups, bool trace_min_groups) {
  TORCH_CHECK(!fields.empty(), "Batch dimension must be non-empty");
  TORCH_CHECK(thresholds.size() == fields.size(), "one persistence threshold per field is required");
  const auto dev = fields[0].device();
  TORCH_CHECK(dev.is_cuda(), "GPU backend requires a CUDA or MPS tensor, got ", dev);
  for (auto& f : fields) {
    validate(f);
    TORCH_CHECK(f.dim() == 2, "each field of a list batch must be 2D [H, W]");
    TORCH_CHECK(f.device() == dev, "all fields of a batch must be on one device");
  }
  c10::hip::HIPGuard guard(dev.index());
  dmsc_context* ctx = context_for(dev.index());
  const c10::hip::HIPStream stream = c10::hip::getCurrentHIPStream(dev.index());
  const uint32_t flags = make_flags(return_gradient, is_dual, trace_max_arcs, trace_min_arcs, trace_max_groups,
                                    trace_min_groups);
  std::vector<torch::Tensor> contig;
  std::vector<const float*> ptrs;
  std::vector<int32_t> hs, ws;
  for (auto& f : fields) {
    contig.push_back(f.contiguous());
    ptrs.push_back(contig.back().data_ptr<float>());
    hs.push_back((int32_t)f.size(0));
    ws.push_back((int32_t)f.size(1));
  }
  std::vector<float> thr(thresholds.begin(), thresholds.end());
  return run_batch(ctx, dev, ptrs, hs, ws, thr, flags, stream);
}

// The C-ABI view of per-field pair tensors (sad_pts, p_min, ppairs_min, min_pts, p_max, ppairs_max,
// max_pts) of an unsimplified complex; contiguous copies are kept alive in `keep`.
std::vector<dmsc_pairs> pair_views(std::vector<std::vector<torch::Tensor>>& pairs, torch::Device dev,
                                   std::vector<torch::Tensor>& keep) {
  std::vector<dmsc_pairs> pv(pairs.size());
  for (size_t b = 0; b < pairs.size(); ++b) {
    auto& p = pairs[b];
    TORCH_CHECK(p.size() == 7, "pairs must hold (sad_pts, p_min, ppairs_min, min_pts, p_max, ppairs_max, max_pts)");
    for (int k = 0; k < 7; ++k) {
      TORCH_CHECK(p[k].device() == dev, "pair tensors must be on the fields' device");
      p[k] = p[k].contiguous();
      keep.push_back(p[k]);
      const bool fl = (k == 1 || k == 4);
      TORCH_CHECK(p[k].scalar_type() == (fl ? torch::kFloat32 : torch::kInt32), "pair tensor ", k, " has the wrong dtype");
    }
    const int64_t ns = p[0].numel();
    TORCH_CHECK(p[1].numel() == ns && p[2].numel() == ns && p[4].numel() == ns && p[5].numel() == ns,
                "per-saddle pair tensors must have one entry per saddle");
    pv[b] = dmsc_pairs{ns,
                       p[3].numel(),
                       p[6].numel(),
                       ns ? p[0].data_ptr<int32_t>() : nullptr,
                       ns ? p[1].data_ptr<float>() : nullptr,
                       ns ? p[2].data_ptr<int32_t>() : nullptr,
                       p[3].numel() ? p[3].data_ptr<int32_t>() : nullptr,
                       ns ? p[4].data_ptr<float>() : nullptr,
                       ns ? p[5].data_ptr<int32_t>() : nullptr,
                       p[6].numel() ? p[6].data_ptr<int32_t>() : nullptr};
  }
  return pv;
}

// Backward of the persistence loss for a list of fields; pairs[b] = the pair tensors of field b's
// unsimplified complex.
std::vector<torch::Tensor> persistence_backward(std::vector<torch::Tensor> fields,
                                                std::vector<std::vector<torch::Tensor>> pairs,
                                                torch::Tensor grad_scale, double max_persistence, bool is_dual) {
  const int64_t B = (int64_t)fields.size();
  TORCH_CHECK(B > 0 && (int64_t)pairs.size() == B, "one set of pairs per field is required");
  const auto dev = fields[0].device();
  TORCH_CHECK(dev.is_cuda(), "persistence_backward requires GPU tensors");
  TORCH_CHECK(grad_scale.device() == dev && grad_scale.scalar_type() == torch::kFloat32 && grad_scale.numel() == B,
              "grad_scale must be a float32 tensor with one entry per field on the fields' device");
  grad_scale = grad_scale.contiguous();
  c10::hip::HIPGuard guard(dev.index());
  dmsc_context* ctx = context_for(dev.index());
  hipStream_t stream = c10::hip::getCurrentHIPStream(dev.index()).stream();
  std::vector<torch::Tensor> keep, grads;
  std::vector<const float*> ptrs;
  std::vector<float*> gptrs;
  std::vector<int32_t> hs, ws;
  for (int64_t b = 0; b < B; ++b) {
    torch::Tensor f = fields[b];
    validate(f);
    TORCH_CHECK(f.dim() == 2 && f.device() == dev, "each field must be a 2D tensor on the same device");
    f = f.contiguous();
    keep.push_back(f);
    ptrs.push_back(f.data_ptr<float>());
    hs.push_back((int32_t)f.size(0));
    ws.push_back((int32_t)f.size(1));
    grads.push_back(torch::empty_like(f));
    gptrs.push_back(grads.back().data_ptr<float>());
  }
  std::vector<dmsc_pairs> pv = pair_views(pairs, dev, keep);
  dmsc_status st;
  {
    pybind11::gil_scoped_release release;
    st = dmsc_persistence_backward(ctx, (int32_t)B, ptrs.data(), hs.data(), ws.data(), is_dual ? DMSC_DUAL : 0u,
                                   pv.data(), (float)max_persistence, grad_scale.data_ptr<float>(), gptrs.data(),
                                   (dmsc_stream_t)stream);
  }
  TORCH_CHECK(st == DMSC_OK, dmsc_last_error(ctx));
  return grads;
}

// Per-field loss sum(pers^2 for 0 <= pers < max_persistence) of the pairs of each field, as a [B]
// float32 tensor on their device (one kernel for all fields).
torch::Tensor persistence_loss(std::vector<std::vector<torch::Tensor>> pairs, double max_persistence) {
  const int64_t B = (int64_t)pairs.size();
  TORCH_CHECK(B > 0 && !pairs[0].empty(), "one set of pairs per field is required");
  const auto dev = pairs[0][0].device();
  TORCH_CHECK(dev.is_cuda(), "persistence_loss requires GPU tensors");
  c10::hip::HIPGuard guard(dev.index());
  dmsc_context* ctx = context_for(dev.index());
  hipStream_t stream = c10::hip::getCurrentHIPStream(dev.index()).stream();
  std::vector<torch::Tensor> keep;
  std::vector<dmsc_pairs> pv = pair_views(pairs, dev, keep);
  torch::Tensor loss = torch::empty({B}, torch::TensorOptions().dtype(torch::kFloat32).device(dev));
  const dmsc_status st = dmsc_persistence_loss(ctx, (int32_t)B, pv.data(), (float)max_persistence,
                                               loss.data_ptr<float>(), (dmsc_stream_t)stream);
  TORCH_CHECK(st == DMSC_OK, dmsc_last_error(ctx));
  return loss;
}

// Element type of each debug_fetch name: the host buffer is sized here, never by the caller, so a
// name cannot write past it.
torch::ScalarType debug_dtype(const std::string& name) {
  if (name == "pers_min" || name == "pers_max") return torch::kFloat32;
  if (name == "arc_stats" || name == "face_pj" || name == "vert_pj") return torch::kInt64;
  return torch::kInt32;
}

pybind11::object debug_fetch(int device, const std::string& name, int64_t capacity) {
  TORCH_CHECK(capacity >= 0, "dmsc: debug_fetch capacity must be >= 0");
  dmsc_context* ctx = context_for(device);
  torch::Tensor t = torch::empty({capacity}, torch::TensorOptions().dtype(debug_dtype(name)));
  int64_t n = dmsc_debug_fetch(ctx, name.c_str(), t.data_ptr(), capacity);
  if (n < 0) return pybind11::none();
  return pybind11::cast(t.slice(0, 0, std::min(n, capacity)));
}

void profile_enable(int device, int level) {
  TORCH_CHECK(dmsc_profile_enable(context_for(device), level) == DMSC_OK, "dmsc: profile level must be 0, 1 or 2");
}

void set_param(int device, const std::string& name, int64_t value) {
  TORCH_CHECK(dmsc_set_param(context_for(device), name.c_str(), value) == DMSC_OK,
              "dmsc: unknown parameter or value out of range: ", name, "=", value);
}
int64_t get_param(int device, const std::string& name) {
  int64_t v = 0;
  TORCH_CHECK(dmsc_get_param(context_for(device), name.c_str(), &v) == DMSC_OK, "dmsc: unknown parameter: ", name);
  return v;
}

pybind11::dict profile_read(int device) {
  dmsc_context* ctx = context_for(device);
  pybind11::dict d;
  for (int32_t i = 0; i < dmsc_profile_count(ctx); ++i) {
    const char* name = nullptr;
    double ms = 0;
    int64_t n = 0;
    if (dmsc_profile_entry(ctx, i, &name, &ms, &n) == DMSC_OK) d[name] = pybind11::make_tuple(ms, n);
  }
  return d;
}

}  // namespace

PYBIND11_MODULE(dmsc_gpu, m) {
  m.def("profile_enable", &profile_enable, pybind11::arg("device"), pybind11::arg("on"));
  m.def("profile_read", &profile_read, pybind11::arg("device"));
  m.def("set_param", &set_param, pybind11::arg("device"), pybind11::arg("name"), pybind11::arg("value"));
  m.def("get_param", &get_param, pybind11::arg("device"), pybind11::arg("name"));
  auto cls = pybind11::class_<DMSComplex>(m, "DMSComplex", pybind11::module_local(),
                                          "Discrete Morse-Smale complex (19 tensors, reference slot order)");
  for (int i = 0; i < DMSC_NUM_OUTPUTS; ++i)
    cls.def_property_readonly(kNames[i], [i](const DMSComplex& r) { return r.t[i]; });
  cls.def("__getitem__", [](const DMSComplex& r, int idx) {
    if (idx < 0 || idx >= DMSC_NUM_OUTPUTS) throw pybind11::index_error();
    return r.t[idx];
  });
  cls.def("__len__", [](const DMSComplex&) { return DMSC_NUM_OUTPUTS; });
  // unpacking (MSComplex(*result)) takes the 19 tensors in one call instead of 19 __getitem__ calls
  // and the IndexError that ends the sequence protocol
  cls.def("__iter__", [](const DMSComplex& r) {
    pybind11::tuple t((size_t)DMSC_NUM_OUTPUTS);
    for (int i = 0; i < DMSC_NUM_OUTPUTS; ++i) t[i] = pybind11::cast(r.t[i]);
    return pybind11::iter(t);
  });

  m.def("extract_dmsc", &extract_dmsc, "MI355X Discrete Morse-Smale complex computation",
        pybind11::arg("scalar_field"), pybind11::arg("persistence_threshold"), pybind11::arg("return_gradient") = false,
        pybind11::arg("is_dual") = false, pybind11::arg("trace_max_arcs") = false,
        pybind11::arg("trace_min_arcs") = false, pybind11::arg("trace_max_groups") = false,
        pybind11::arg("trace_min_groups") = false);
  m.def("extract_dmsc_list", &extract_dmsc_list, "Mixed-size batch of fields with per-field thresholds",
        pybind11::arg("fields"), pybind11::arg("persistence_thresholds"), pybind11::arg("return_gradient") = false,
        pybind11::arg("is_dual") = false, pybind11::arg("trace_max_arcs") = false,
        pybind11::arg("trace_min_arcs") = false, pybind11::arg("trace_max_groups") = false,
        pybind11::arg("trace_min_groups") = false);
  m.def("persistence_backward", &persistence_backward, "Gradient of sum(pers^2 for 0 <= pers < max) per field",
        pybind11::arg("fields"), pybind11::arg("pairs"), pybind11::arg("grad_scale"),
        pybind11::arg("max_persistence"), pybind11::arg("is_dual") = false);
  m.def("persistence_loss", &persistence_loss, "sum(pers^2 for 0 <= pers < max) of each field's pairs",
        pybind11::arg("pairs"), pybind11::arg("max_persistence"));
  m.def("debug_fetch", &debug_fetch, pybind11::arg("device"), pybind11::arg("name"), pybind11::arg("capacity"));
}
