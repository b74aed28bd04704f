// bindings/pymodule.cpp -- pybind11 module `_usx` over the GPU-backed host API.
//
// Class, attribute and argument names follow the reference's
// bindings/pymodule.cpp:53-622 for everything on the imaging path, the
// simulator that feeds it, and the display helpers, so Python code written
// against the reference runs unchanged. Array accessors are zero-copy views
// tied to their owner (reference :32-42); work-heavy calls release the GIL.
// Doppler and elastography follow reference :486-545; scan conversion is not
// part of this module.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <algorithm>

#include "../core/src/gpu.hpp"
#include "usx/beamformer.hpp"
#include "usx/demodulate.hpp"
#include "usx/doppler.hpp"
#include "usx/geometry.hpp"
#include "usx/pipeline.hpp"
#include "usx/simulator.hpp"
#include "usx/types.hpp"
#include "usx_gpu.h"

namespace py = pybind11;
using namespace usx;

namespace {

template <typename T>
py::array_t<T> view(T* ptr, std::vector<py::ssize_t> shape, py::object owner) {
  std::vector<py::ssize_t> strides(shape.size());
  py::ssize_t s = static_cast<py::ssize_t>(sizeof(T));
  for (std::size_t i = shape.size(); i-- > 0;) {
    strides[i] = s;
    s *= shape[i];
  }
  return py::array_t<T>(shape, strides, ptr, std::move(owner));
}

template <typename T>
py::array_t<T> copy_vector(const std::vector<T>& v, std::vector<py::ssize_t> shape) {
  py::array_t<T> a(shape);
  std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(T));
  return a;
}

// Device memory for callers without an allocator of their own (tests and
// graph-style callers of DeviceRunner): usx_buffer_alloc on the calling
// thread's context, freed when the object goes; copies are synchronous.
class DeviceBuffer {
 public:
  explicit DeviceBuffer(std::size_t bytes) : ctx_(thread_context().get()), n_(bytes) {
    check(ctx_, usx_buffer_alloc(ctx_, bytes, &p_), "buffer_alloc");
  }
  ~DeviceBuffer() {
    if (p_) usx_buffer_free(ctx_, p_);
  }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  std::uintptr_t ptr() const { return reinterpret_cast<std::uintptr_t>(p_); }
  std::size_t nbytes() const { return n_; }
  void upload(const py::array& a) {
    const py::array c = py::array::ensure(a, py::array::c_style);
    if (static_cast<std::size_t>(c.nbytes()) > n_)
      throw std::invalid_argument("DeviceBuffer.upload: array larger than the buffer");
    check(ctx_, usx_buffer_upload(ctx_, p_, c.data(), static_cast<std::size_t>(c.nbytes())),
          "buffer_upload");
  }
  py::array download(const py::object& dtype, const std::vector<py::ssize_t>& shape) {
    py::array a(py::dtype::from_args(dtype), shape);
    if (static_cast<std::size_t>(a.nbytes()) > n_)
      throw std::invalid_argument("DeviceBuffer.download: shape larger than the buffer");
    check(ctx_, usx_buffer_download(ctx_, a.mutable_data(), p_, static_cast<std::size_t>(a.nbytes())),
          "buffer_download");
    return a;
  }
  void zero() { check(ctx_, usx_buffer_zero(ctx_, p_, n_), "buffer_zero"); check(ctx_, usx_ctx_synchronize(ctx_), "synchronize"); }

 private:
  usx_ctx* ctx_;
  void* p_ = nullptr;
  std::size_t n_;
};

// Device-resident runner for benchmarks and graph-style callers: descriptors
// are built once from a template ChannelData, then frames are beamformed
// from device pointers (e.g. torch tensors' data_ptr()) on the runner's own
// stream, asynchronously. Stage/kernel timings come from HIP events.
class DeviceRunner {
 public:
  DeviceRunner(const ScanGrid& grid, const PipelineConfig& cfg, const ChannelData& like)
      : grid_(grid), cfg_(cfg) {
    like.validate();
    grid_.validate();
    cfg_.beamformer.validate();
    is_rf_ = !like.is_iq;
    if (is_rf_) {
      ChannelData tmp = like;
      tmp.rf.clear();
      dplan_ = make_demod_plan(tmp, cfg_.demod);
      bplan_ = make_bf_plan(like, dplan_.n_out, dplan_.fs_out, true, dplan_.fdem, grid_,
                            cfg_.beamformer, 0, like.n_events());
    } else {
      bplan_ = make_bf_plan(like, like.n_samples(), like.acq.sampling_frequency, true,
                            like.demodulation_frequency, grid_, cfg_.beamformer, 0,
                            like.n_events());
    }
    // Re-point the descriptor tables at this object's copies.
    dplan_.desc.taps = dplan_.taps.data();
    dplan_.desc.rotator = reinterpret_cast<const usx_c32*>(dplan_.rotator.data());
    fix_bf();
  }
  ~DeviceRunner() { drop_graph(); }
  DeviceRunner(const DeviceRunner&) = delete;
  DeviceRunner& operator=(const DeviceRunner&) = delete;
  void fix_bf() {
    bplan_.desc.element_x = bplan_.element_x.data();
    bplan_.desc.tx = bplan_.tx.data();
    bplan_.desc.axis0 = bplan_.axis0.data();
    bplan_.desc.axis1 = bplan_.axis1.data();
    bplan_.desc.axis0_sin = bplan_.s0.empty() ? nullptr : bplan_.s0.data();
    bplan_.desc.axis0_cos = bplan_.c0.empty() ? nullptr : bplan_.c0.data();
  }
  void run(std::uintptr_t in_ptr, std::uintptr_t frame_ptr) {
    usx_ctx* c = ctx_.get();
    if (is_rf_) {
      check(c, usx_process_frame(c, &dplan_.desc, &bplan_.desc, reinterpret_cast<const float*>(in_ptr),
                                 reinterpret_cast<usx_c32*>(frame_ptr), USX_DEVICE, nullptr),
            "process_frame");
    } else {
      check(c, usx_beamform(c, &bplan_.desc, reinterpret_cast<const void*>(in_ptr),
                            reinterpret_cast<usx_c32*>(frame_ptr), USX_DEVICE),
            "beamform");
    }
  }
  // run() through the allocation-free entry point (RF runners only): the
  // scratch comes from the library's fixed per-GPU array, not the context.
  void run_noalloc(std::uintptr_t in_ptr, std::uintptr_t frame_ptr) {
    if (!is_rf_) throw std::invalid_argument("run_noalloc: the runner's template must be RF");
    usx_ctx* c = ctx_.get();
    check(c, usx_process_frame_noalloc(c, &dplan_.desc, &bplan_.desc,
                                       reinterpret_cast<const float*>(in_ptr),
                                       reinterpret_cast<usx_c32*>(frame_ptr)),
          "process_frame_noalloc");
  }
  // The frame of run() with the RF and frame pointers bound now, recorded
  // into a HIP graph (a one-frame usx_plan: one eager pass, then the
  // capture); run_graph() replays it on the runner's stream.
  void capture(std::uintptr_t in_ptr, std::uintptr_t frame_ptr) {
    if (!is_rf_) throw std::invalid_argument("capture: the runner's template must be RF");
    usx_ctx* c = ctx_.get();
    if (graph_) usx_plan_destroy(c, graph_);
    graph_ = nullptr;
    // A capture on the caller's buffers ends an earlier capture_host: its
    // device copies go, so graph_frame() cannot return that capture's frame.
    if (graph_rf_ && reinterpret_cast<void*>(in_ptr) != graph_rf_) {
      usx_buffer_free(c, graph_rf_);
      graph_rf_ = nullptr;
    }
    if (graph_frame_ && reinterpret_cast<void*>(frame_ptr) != graph_frame_) {
      usx_buffer_free(c, graph_frame_);
      graph_frame_ = nullptr;
    }
    const float* rf = reinterpret_cast<const float*>(in_ptr);
    usx_c32* fr = reinterpret_cast<usx_c32*>(frame_ptr);
    check(c, usx_plan_create(c, 1, &dplan_.desc, &bplan_.desc, &rf, &fr, &graph_), "plan_create");
  }
  void run_graph() {
    if (!graph_) throw std::logic_error("run_graph: no captured frame (call capture first)");
    usx_ctx* c = ctx_.get();
    check(c, usx_plan_launch(c, graph_, 1), "plan_launch");
  }
  // capture() on device copies of host channel data of the template's shape,
  // which the runner keeps with the graph; graph_frame() downloads the
  // frame the replays write.
  void capture_host(const ChannelData& data) {
    if (!is_rf_) throw std::invalid_argument("capture_host: the runner's template must be RF");
    if (data.rf.size() != static_cast<std::size_t>(dplan_.desc.n_traces) * dplan_.desc.n_in)
      throw std::invalid_argument("channel data does not match the runner's template");
    usx_ctx* c = ctx_.get();
    drop_graph();
    check(c, usx_buffer_alloc(c, data.rf.size() * sizeof(float), &graph_rf_), "buffer_alloc");
    check(c, usx_buffer_alloc(c, sizeof(usx_c32) * frame_pixels(), &graph_frame_), "buffer_alloc");
    check(c, usx_buffer_upload(c, graph_rf_, data.rf.data(), data.rf.size() * sizeof(float)),
          "buffer_upload");
    capture(reinterpret_cast<std::uintptr_t>(graph_rf_), reinterpret_cast<std::uintptr_t>(graph_frame_));
  }
  py::array_t<std::complex<float>> graph_frame() {
    if (!graph_frame_) throw std::logic_error("graph_frame: no frame buffer (call capture_host first)");
    py::array_t<std::complex<float>> a({bplan_.desc.n_axis0, bplan_.desc.n_axis1});
    usx_ctx* c = ctx_.get();
    check(c, usx_buffer_download(c, a.mutable_data(), graph_frame_, sizeof(usx_c32) * frame_pixels()),
          "buffer_download");
    return a;
  }
  void drop_graph() {
    usx_ctx* c = ctx_.get();
    if (graph_) usx_plan_destroy(c, graph_);
    if (graph_rf_) usx_buffer_free(c, graph_rf_);
    if (graph_frame_) usx_buffer_free(c, graph_frame_);
    graph_ = nullptr;
    graph_rf_ = graph_frame_ = nullptr;
  }
  void synchronize() {
    usx_ctx* c = ctx_.get();
    check(c, usx_ctx_synchronize(c), "synchronize");
  }
  std::uintptr_t stream() { return reinterpret_cast<std::uintptr_t>(usx_ctx_stream(ctx_.get())); }
  void record(int slot) {
    usx_ctx* c = ctx_.get();
    check(c, usx_ctx_event_record(c, slot), "event_record");
  }
  float elapsed(int a, int b) {
    usx_ctx* c = ctx_.get();
    float ms = 0;
    check(c, usx_ctx_event_elapsed(c, a, b, &ms), "event_elapsed");
    return ms;
  }
  void profile(bool on) {
    usx_ctx* c = ctx_.get();
    check(c, usx_ctx_profile(c, on ? 1 : 0), "profile");
  }
  std::string profile_report() {
    usx_ctx* c = ctx_.get();
    std::vector<char> buf(1 << 16);
    check(c, usx_ctx_profile_report(c, buf.data(), buf.size()), "profile_report");
    return std::string(buf.data());
  }
  std::vector<py::ssize_t> input_shape() const {
    const auto& b = bplan_.desc;
    if (is_rf_) return {b.n_events, b.n_channels, dplan_.desc.n_in};
    return {b.n_events, b.n_channels, b.n_samples};
  }
  bool is_rf() const { return is_rf_; }
  int iq_samples() const { return bplan_.desc.n_samples; }
  int frame_pixels() const { return bplan_.desc.n_axis0 * bplan_.desc.n_axis1; }

  // Pick the launch shape of this runner's frame by timing the candidates
  // (usx_autotune) on the given device buffers; returns (slots, mv
  // occupancy, ms per frame) and keeps the choice for run().
  py::tuple autotune(std::uintptr_t in_ptr, std::uintptr_t frame_ptr, int reps) {
    if (!is_rf_) throw std::invalid_argument("autotune: the runner's template must be RF");
    usx_ctx* c = ctx_.get();
    float ms = 0.0f;
    {
      py::gil_scoped_release release;
      check(c, usx_autotune(c, &dplan_.desc, &bplan_.desc, reinterpret_cast<const float*>(in_ptr),
                            reinterpret_cast<usx_c32*>(frame_ptr), reps, &ms),
            "autotune");
    }
    return py::make_tuple(bplan_.desc.tile_slots, bplan_.desc.tile_mv_occupancy, ms);
  }

  // The same on host channel data of the template's shape, uploaded into
  // temporary device buffers.
  py::tuple autotune_host(const ChannelData& data, int reps) {
    if (!is_rf_) throw std::invalid_argument("autotune: the runner's template must be RF");
    if (data.rf.size() != static_cast<std::size_t>(dplan_.desc.n_traces) * dplan_.desc.n_in)
      throw std::invalid_argument("channel data does not match the runner's template");
    usx_ctx* c = ctx_.get();
    void* rf = nullptr;
    void* fr = nullptr;
    const std::size_t nf = static_cast<std::size_t>(bplan_.desc.n_axis0) * bplan_.desc.n_axis1;
    check(c, usx_buffer_alloc(c, data.rf.size() * sizeof(float), &rf), "buffer_alloc");
    const usx_status st = usx_buffer_alloc(c, nf * sizeof(usx_c32), &fr);
    if (st != USX_OK) {
      usx_buffer_free(c, rf);
      check(c, st, "buffer_alloc");
    }
    try {
      check(c, usx_buffer_upload(c, rf, data.rf.data(), data.rf.size() * sizeof(float)), "buffer_upload");
      py::tuple r = autotune(reinterpret_cast<std::uintptr_t>(rf), reinterpret_cast<std::uintptr_t>(fr), reps);
      usx_buffer_free(c, rf);
      usx_buffer_free(c, fr);
      return r;
    } catch (...) {
      usx_buffer_free(c, rf);
      usx_buffer_free(c, fr);
      throw;
    }
  }

  // Device-resident ms per frame on host channel data of the template's
  // shape: uploaded once into temporary device buffers, one untimed frame,
  // then `reps` frames between two events on the context's stream.
  float time_host(const ChannelData& data, int reps) {
    if (!is_rf_) throw std::invalid_argument("time_host: the runner's template must be RF");
    if (data.rf.size() != static_cast<std::size_t>(dplan_.desc.n_traces) * dplan_.desc.n_in)
      throw std::invalid_argument("channel data does not match the runner's template");
    if (reps < 1) throw std::invalid_argument("time_host: reps must be >= 1");
    usx_ctx* c = ctx_.get();
    void* rf = nullptr;
    void* fr = nullptr;
    const std::size_t nf = static_cast<std::size_t>(bplan_.desc.n_axis0) * bplan_.desc.n_axis1;
    check(c, usx_buffer_alloc(c, data.rf.size() * sizeof(float), &rf), "buffer_alloc");
    const usx_status st = usx_buffer_alloc(c, nf * sizeof(usx_c32), &fr);
    if (st != USX_OK) {
      usx_buffer_free(c, rf);
      check(c, st, "buffer_alloc");
    }
    try {
      check(c, usx_buffer_upload(c, rf, data.rf.data(), data.rf.size() * sizeof(float)), "buffer_upload");
      const auto in = reinterpret_cast<std::uintptr_t>(rf), out = reinterpret_cast<std::uintptr_t>(fr);
      run(in, out);
      record(0);
      for (int i = 0; i < reps; ++i) run(in, out);
      record(1);
      synchronize();
      const float ms = elapsed(0, 1) / static_cast<float>(reps);
      usx_buffer_free(c, rf);
      usx_buffer_free(c, fr);
      return ms;
    } catch (...) {
      usx_buffer_free(c, rf);
      usx_buffer_free(c, fr);
      throw;
    }
  }

  // ---- learned apodization (no reference counterpart) ----------------------
  using farr = py::array_t<float, py::array::c_style | py::array::forcecast>;

  void set_apodization(farr p) {
    if (p.size() != USX_APOD_PARAMS)
      throw std::invalid_argument("set_apodization: expected " + std::to_string(USX_APOD_PARAMS) +
                                  " parameters");
    usx_ctx* c = ctx_.get();
    check(c, usx_apod_set_params(c, p.data()), "apod_set_params");
  }
  py::array_t<float> get_apodization() {
    py::array_t<float> p(USX_APOD_PARAMS);
    usx_ctx* c = ctx_.get();
    check(c, usx_apod_get_params(c, p.mutable_data()), "apod_get_params");
    return p;
  }
  // The descriptor with the minimum-variance combiner (the labels' source).
  usx_bf_desc mv_desc() const {
    usx_bf_desc b = bplan_.desc;
    b.combiner = 2;
    return b;
  }
  // Host ChannelData (the template's shape) -> the IQ the plan beamforms.
  std::vector<usx_c32> host_iq(const ChannelData& data) {
    usx_ctx* c = ctx_.get();
    const auto& b = bplan_.desc;
    std::vector<usx_c32> iq(static_cast<std::size_t>(b.n_events) * b.n_channels * b.n_samples);
    if (is_rf_) {
      if (data.rf.size() != static_cast<std::size_t>(dplan_.desc.n_traces) * dplan_.desc.n_in)
        throw std::invalid_argument("channel data does not match the runner's template");
      check(c, usx_demodulate(c, &dplan_.desc, data.rf.data(), iq.data(), USX_HOST), "demodulate");
    } else {
      if (data.iq.size() != iq.size())
        throw std::invalid_argument("channel data does not match the runner's template");
      std::memcpy(iq.data(), data.iq.data(), iq.size() * sizeof(usx_c32));
    }
    return iq;
  }
  py::tuple mv_weights(const ChannelData& data) {
    const auto& b = bplan_.desc;
    const std::vector<usx_c32> iq = host_iq(data);
    py::array_t<std::complex<float>> frame({b.n_axis0, b.n_axis1});
    py::array_t<std::complex<float>> w({b.n_axis0, b.n_axis1, b.n_channels});
    py::array_t<int> st({b.n_axis0, b.n_axis1});
    const usx_bf_desc d = mv_desc();
    usx_ctx* c = ctx_.get();
    {
      py::gil_scoped_release release;
      check(c, usx_mv_weights(c, &d, iq.data(), reinterpret_cast<usx_c32*>(frame.mutable_data()),
                              reinterpret_cast<usx_c32*>(w.mutable_data()), st.mutable_data(), USX_HOST),
            "mv_weights");
    }
    return py::make_tuple(frame, w, st);
  }
  void dataset_reserve(std::size_t n) {
    usx_ctx* c = ctx_.get();
    check(c, usx_apod_dataset_reserve(c, n), "apod_dataset_reserve");
  }
  std::size_t dataset_size() {
    usx_ctx* c = ctx_.get();
    std::size_t n = 0;
    check(c, usx_apod_dataset_size(c, &n), "apod_dataset_size");
    return n;
  }
  py::array_t<std::complex<float>> dataset_append(
      const ChannelData& data, py::array_t<int, py::array::c_style | py::array::forcecast> pixels) {
    const auto& b = bplan_.desc;
    py::array_t<std::complex<float>> frame({b.n_axis0, b.n_axis1});
    const usx_bf_desc d = mv_desc();
    usx_ctx* c = ctx_.get();
    const void* src;
    if (is_rf_) {
      if (data.rf.size() != static_cast<std::size_t>(dplan_.desc.n_traces) * dplan_.desc.n_in)
        throw std::invalid_argument("channel data does not match the runner's template");
      src = data.rf.data();
    } else {
      if (data.iq.size() != static_cast<std::size_t>(b.n_events) * b.n_channels * b.n_samples)
        throw std::invalid_argument("channel data does not match the runner's template");
      src = data.iq.data();
    }
    {
      py::gil_scoped_release release;
      check(c, usx_apod_dataset_append(c, is_rf_ ? &dplan_.desc : nullptr, &d, src, USX_HOST,
                                       pixels.data(), static_cast<int>(pixels.size()),
                                       reinterpret_cast<usx_c32*>(frame.mutable_data())),
            "apod_dataset_append");
    }
    return frame;
  }
  py::array_t<float> train(int n_steps, int batch, float lr, float beta1, float beta2, float eps,
                           float weight_decay, unsigned long long seed) {
    usx_apod_train_desc t{};
    t.n_steps = n_steps;
    t.batch = batch;
    t.lr = lr;
    t.beta1 = beta1;
    t.beta2 = beta2;
    t.eps = eps;
    t.weight_decay = weight_decay;
    t.seed = seed;
    py::array_t<float> losses(std::max(n_steps, 0));
    usx_ctx* c = ctx_.get();
    {
      py::gil_scoped_release release;
      check(c, usx_apod_train(c, &t, losses.mutable_data()), "apod_train");
    }
    return losses;
  }
  py::tuple gradient(py::array_t<uint16_t, py::array::c_style | py::array::forcecast> x, farr g,
                     py::array_t<uint32_t, py::array::c_style | py::array::forcecast> mask,
                     farr sw) {
    const py::ssize_t n = sw.size();
    if (x.size() != n * USX_APOD_FEATURES || g.size() != n * USX_APOD_FEATURES || mask.size() != n * 4)
      throw std::invalid_argument("gradient: x, g (n x 256), mask (n x 4) and sw (n) disagree");
    float loss = 0.0f;
    py::array_t<float> grads(USX_APOD_PARAMS);
    usx_ctx* c = ctx_.get();
    check(c, usx_apod_gradient(c, static_cast<int>(n), x.data(), g.data(), mask.data(), sw.data(),
                               &loss, grads.mutable_data()),
          "apod_gradient");
    return py::make_tuple(loss, grads);
  }

 private:
  ScanGrid grid_;
  PipelineConfig cfg_;
  bool is_rf_ = true;
  DemodPlan dplan_;
  BeamformPlan bplan_;
  GpuContext ctx_;
  usx_plan* graph_ = nullptr;
  void* graph_rf_ = nullptr;     // device copies owned for capture_host
  void* graph_frame_ = nullptr;
};

// A fixed sequence of RF frames, each with its own probe, grid, transmits and
// configuration, recorded once into a HIP graph (usx_plan_create) and
// replayed. The sequence owns the device memory: the RF of every distinct
// ChannelData (frames that share one are beamformed from one copy) and one
// output frame per entry.
class FrameSequence {
 public:
  FrameSequence(const std::vector<ChannelData*>& likes, const std::vector<ScanGrid*>& grids,
                const std::vector<PipelineConfig>& cfgs) {
    const std::size_t n = likes.size();
    if (n == 0 || grids.size() != n || cfgs.size() != n)
      throw std::invalid_argument("FrameSequence: likes, grids and configs must have one entry per frame");
    usx_ctx* c = ctx_.get();
    dplans_.resize(n);
    bplans_.resize(n);
    std::vector<const float*> rf(n);
    std::vector<usx_c32*> fr(n);
    for (std::size_t f = 0; f < n; ++f) {
      const ChannelData& like = *likes[f];
      like.validate();
      grids[f]->validate();
      cfgs[f].beamformer.validate();
      if (like.is_iq || like.rf.empty()) throw std::invalid_argument("FrameSequence: frames must be RF");
      ChannelData tmp;
      tmp.probe = like.probe;
      tmp.medium = like.medium;
      tmp.acq = like.acq;
      tmp.transmits = like.transmits;
      tmp.is_iq = false;
      dplans_[f] = make_demod_plan(tmp, cfgs[f].demod);
      bplans_[f] = make_bf_plan(like, dplans_[f].n_out, dplans_[f].fs_out, true, dplans_[f].fdem,
                                *grids[f], cfgs[f].beamformer, 0, like.n_events());
      auto it = rf_of_.find(likes[f]);
      if (it == rf_of_.end()) {
        void* p = alloc(like.rf.size() * sizeof(float));
        check(c, usx_buffer_upload(c, p, like.rf.data(), like.rf.size() * sizeof(float)), "buffer_upload");
        it = rf_of_.emplace(likes[f], static_cast<const float*>(p)).first;
      }
      rf[f] = it->second;
      shapes_.push_back({grids[f]->n_axis0, grids[f]->n_axis1});
      fr[f] = static_cast<usx_c32*>(alloc(sizeof(usx_c32) * grids[f]->n_pixels()));
      frames_.push_back(fr[f]);
    }
    std::vector<usx_demod_desc> d(n);
    std::vector<usx_bf_desc> b(n);
    for (std::size_t f = 0; f < n; ++f) {
      DemodPlan& dp = dplans_[f];
      BeamformPlan& bp = bplans_[f];
      dp.desc.taps = dp.taps.data();
      dp.desc.rotator = reinterpret_cast<const usx_c32*>(dp.rotator.data());
      bp.desc.element_x = bp.element_x.data();
      bp.desc.tx = bp.tx.data();
      bp.desc.axis0 = bp.axis0.data();
      bp.desc.axis1 = bp.axis1.data();
      bp.desc.axis0_sin = bp.s0.empty() ? nullptr : bp.s0.data();
      bp.desc.axis0_cos = bp.c0.empty() ? nullptr : bp.c0.data();
      d[f] = dp.desc;
      b[f] = bp.desc;
    }
    py::gil_scoped_release release;
    check(c, usx_plan_create(c, static_cast<int>(n), d.data(), b.data(), rf.data(), fr.data(), &plan_),
          "plan_create");
  }
  ~FrameSequence() {
    usx_ctx* c = ctx_.get();
    if (plan_) usx_plan_destroy(c, plan_);
    for (void* p : owned_) usx_buffer_free(c, p);
  }
  FrameSequence(const FrameSequence&) = delete;
  FrameSequence& operator=(const FrameSequence&) = delete;

  void launch(bool graph) {
    usx_ctx* c = ctx_.get();
    check(c, usx_plan_launch(c, plan_, graph ? 1 : 0), "plan_launch");
  }
  void synchronize() {
    usx_ctx* c = ctx_.get();
    check(c, usx_ctx_synchronize(c), "synchronize");
  }
  void record(int slot) {
    usx_ctx* c = ctx_.get();
    check(c, usx_ctx_event_record(c, slot), "event_record");
  }
  float elapsed(int a, int b) {
    usx_ctx* c = ctx_.get();
    float ms = 0;
    check(c, usx_ctx_event_elapsed(c, a, b, &ms), "event_elapsed");
    return ms;
  }
  // Zero every output frame (a fill kernel on the sequence's stream).
  void clear() {
    usx_ctx* c = ctx_.get();
    for (std::size_t f = 0; f < frames_.size(); ++f)
      check(c, usx_buffer_zero(c, frames_[f], sizeof(usx_c32) * shapes_[f].first * shapes_[f].second),
            "buffer_zero");
  }
  py::list frames() {
    usx_ctx* c = ctx_.get();
    py::list out;
    for (std::size_t f = 0; f < frames_.size(); ++f) {
      py::array_t<std::complex<float>> a({shapes_[f].first, shapes_[f].second});
      check(c, usx_buffer_download(c, a.mutable_data(), frames_[f],
                                   sizeof(usx_c32) * shapes_[f].first * shapes_[f].second),
            "buffer_download");
      out.append(a);
    }
    return out;
  }
  // Launch shape of each entry as planned: (slots, mv occupancy).
  py::list tiles() const {
    py::list out;
    for (const auto& bp : bplans_) out.append(py::make_tuple(bp.desc.tile_slots, bp.desc.tile_mv_occupancy));
    return out;
  }
  std::size_t size() const { return dplans_.size(); }

 private:
  void* alloc(std::size_t bytes) {
    usx_ctx* c = ctx_.get();
    void* p = nullptr;
    check(c, usx_buffer_alloc(c, bytes, &p), "buffer_alloc");
    owned_.push_back(p);
    return p;
  }
  std::vector<DemodPlan> dplans_;
  std::vector<BeamformPlan> bplans_;
  std::map<const ChannelData*, const float*> rf_of_;
  std::vector<usx_c32*> frames_;
  std::vector<std::pair<py::ssize_t, py::ssize_t>> shapes_;
  std::vector<void*> owned_;
  GpuContext ctx_;
  usx_plan* plan_ = nullptr;
};

}  // namespace

// Device-resident Doppler for benchmarks and device callers: a frame-major
// ensemble already in device memory (e.g. a torch tensor's data_ptr()),
// outputs to device pointers, enqueued on the runner's own stream.
class DopplerRunner {
 public:
  DopplerRunner(const ScanGrid& grid, int n_frames, real prf, real speed_of_sound, real fdem,
                const DopplerConfig& cfg)
      : grid_(grid), n_frames_(n_frames) {
    grid_.validate();
    if (n_frames < 2) throw std::invalid_argument("color_flow: need at least 2 frames in the ensemble");
    if (cfg.wall_filter != WallFilter::None)
      reject_ = polynomial_reject_matrix(
          n_frames, cfg.wall_filter == WallFilter::MeanSubtraction ? 0 : cfg.polynomial_order);
    e_.n_frames = n_frames;
    e_.n_axis0 = grid_.n_axis0;
    e_.n_axis1 = grid_.n_axis1;
    e_.reject = reject_.empty() ? nullptr : reject_.data();
    f_.prf = prf;
    f_.speed_of_sound = speed_of_sound;
    f_.demodulation_frequency = fdem;
    f_.dz = grid_.n_axis1 > 1 ? std::abs(grid_.axis1[1] - grid_.axis1[0]) : 1e-4f;
    f_.loupas = cfg.loupas ? 1 : 0;
    f_.kernel_axial = cfg.kernel_axial;
    f_.kernel_lateral = cfg.kernel_lateral;
    f_.coherence_threshold = cfg.coherence_threshold;
    f_.power_factor = std::pow(10.0f, cfg.power_threshold_db / 10.0f);
  }
  void color_flow(std::uintptr_t ens, std::uintptr_t vel, std::uintptr_t pw, std::uintptr_t var,
                  std::uintptr_t valid) {
    usx_ctx* c = ctx_.get();
    check(c, usx_color_flow(c, &e_, &f_, reinterpret_cast<const usx_c32*>(ens),
                            reinterpret_cast<float*>(vel), reinterpret_cast<float*>(pw),
                            reinterpret_cast<float*>(var), reinterpret_cast<std::uint8_t*>(valid),
                            USX_DEVICE),
          "color_flow");
  }
  void power_doppler(std::uintptr_t ens, std::uintptr_t pw) {
    usx_ctx* c = ctx_.get();
    check(c, usx_power_doppler(c, &e_, reinterpret_cast<const usx_c32*>(ens),
                               reinterpret_cast<float*>(pw), USX_DEVICE),
          "power_doppler");
  }
  void wall_filter(std::uintptr_t ens, std::uintptr_t out) {
    usx_ctx* c = ctx_.get();
    check(c, usx_wall_filter(c, &e_, reinterpret_cast<const usx_c32*>(ens),
                             reinterpret_cast<usx_c32*>(out), USX_DEVICE),
          "apply_wall_filter");
  }
  void synchronize() {
    usx_ctx* c = ctx_.get();
    check(c, usx_ctx_synchronize(c), "synchronize");
  }
  void record(int slot) {
    usx_ctx* c = ctx_.get();
    check(c, usx_ctx_event_record(c, slot), "event_record");
  }
  float elapsed(int a, int b) {
    usx_ctx* c = ctx_.get();
    float ms = 0;
    check(c, usx_ctx_event_elapsed(c, a, b, &ms), "event_elapsed");
    return ms;
  }
  void profile(bool on) {
    usx_ctx* c = ctx_.get();
    check(c, usx_ctx_profile(c, on ? 1 : 0), "profile");
  }
  std::string profile_report() {
    usx_ctx* c = ctx_.get();
    std::vector<char> buf(1 << 16);
    check(c, usx_ctx_profile_report(c, buf.data(), buf.size()), "profile_report");
    return std::string(buf.data());
  }
  int n_frames() const { return n_frames_; }

 private:
  ScanGrid grid_;
  int n_frames_;
  std::vector<real> reject_;
  usx_ensemble_desc e_{};
  usx_color_flow_desc f_{};
  GpuContext ctx_;
};

PYBIND11_MODULE(_usx, m) {
  m.doc() = "usx -- ultrasound imaging path on MI355X (gfx950 HIP kernels behind a C ABI)";

  py::enum_<Window>(m, "Window")
      .value("RECT", Window::Rect)
      .value("HANN", Window::Hann)
      .value("HAMMING", Window::Hamming)
      .value("TUKEY", Window::Tukey)
      .value("BLACKMAN", Window::Blackman);
  py::enum_<TransmitType>(m, "TransmitType")
      .value("FOCUSED", TransmitType::Focused)
      .value("PLANE_WAVE", TransmitType::PlaneWave)
      .value("DIVERGING", TransmitType::Diverging);
  py::enum_<GridType>(m, "GridType")
      .value("CARTESIAN", GridType::Cartesian)
      .value("SECTOR", GridType::Sector);
  py::enum_<Combiner>(m, "Combiner")
      .value("DAS", Combiner::DAS)
      .value("DMAS", Combiner::DMAS)
      .value("MV", Combiner::MV)
      .value("Learned", Combiner::Learned);
  py::enum_<Compounding>(m, "Compounding")
      .value("COHERENT", Compounding::Coherent)
      .value("INCOHERENT", Compounding::Incoherent)
      .value("SELECT", Compounding::Select);
  py::enum_<MvPrecision>(m, "MvPrecision")
      .value("FP32", MvPrecision::FP32)
      .value("BF16", MvPrecision::BF16)
      .value("FP16", MvPrecision::FP16);
  py::enum_<WallFilter>(m, "WallFilter")
      .value("NONE", WallFilter::None)
      .value("POLYNOMIAL", WallFilter::PolynomialRegression)
      .value("MEAN", WallFilter::MeanSubtraction);

  py::class_<Vec3>(m, "Vec3")
      .def(py::init<>())
      .def(py::init<real, real, real>(), py::arg("x"), py::arg("y"), py::arg("z"))
      .def_readwrite("x", &Vec3::x)
      .def_readwrite("y", &Vec3::y)
      .def_readwrite("z", &Vec3::z)
      .def("__repr__", [](const Vec3& v) {
        return "Vec3(" + std::to_string(v.x) + ", " + std::to_string(v.y) + ", " +
               std::to_string(v.z) + ")";
      });

  py::class_<Probe>(m, "Probe")
      .def(py::init<>())
      .def_readwrite("n_elements", &Probe::n_elements)
      .def_readwrite("pitch", &Probe::pitch)
      .def_readwrite("element_width", &Probe::element_width)
      .def_readwrite("elevation_height", &Probe::elevation_height)
      .def_readwrite("center_frequency", &Probe::center_frequency)
      .def_readwrite("bandwidth", &Probe::bandwidth)
      .def("element_position", &Probe::element_position)
      .def("aperture_width", &Probe::aperture_width)
      .def("wavelength", &Probe::wavelength, py::arg("speed_of_sound") = 1540.0f)
      .def_property_readonly("element_x",
                             [](const Probe& p) {
                               std::vector<real> x(static_cast<std::size_t>(p.n_elements));
                               for (int i = 0; i < p.n_elements; ++i) x[i] = p.element_position(i).x;
                               return copy_vector(x, {p.n_elements});
                             })
      .def("validate", &Probe::validate);

  py::class_<Medium>(m, "Medium")
      .def(py::init<>())
      .def_readwrite("speed_of_sound", &Medium::speed_of_sound)
      .def_readwrite("attenuation", &Medium::attenuation)
      .def_readwrite("density", &Medium::density);

  py::class_<Acquisition>(m, "Acquisition")
      .def(py::init<>())
      .def_readwrite("sampling_frequency", &Acquisition::sampling_frequency)
      .def_readwrite("n_samples", &Acquisition::n_samples)
      .def_readwrite("t0", &Acquisition::t0)
      .def_readwrite("prf", &Acquisition::prf)
      .def("max_depth", &Acquisition::max_depth, py::arg("speed_of_sound") = 1540.0f);

  py::class_<Transmit>(m, "Transmit")
      .def(py::init<>())
      .def_readwrite("type", &Transmit::type)
      .def_readwrite("angle", &Transmit::angle)
      .def_readwrite("focus", &Transmit::focus)
      .def_readwrite("origin", &Transmit::origin)
      .def_readwrite("delays", &Transmit::delays)
      .def_readwrite("apodization", &Transmit::apodization)
      .def_property_readonly("n_active", &Transmit::n_active);

  m.def(
      "compute_transmit_delays",
      [](const Probe& probe, const Medium& medium, Transmit& tx, Window w, real f_number) {
        real t0 = 0;
        compute_transmit_delays(probe, medium, tx, w, f_number, &t0);
        return t0;
      },
      py::arg("probe"), py::arg("medium"), py::arg("transmit"), py::arg("window") = Window::Hann,
      py::arg("f_number") = 2.0f);
  m.def("transmit_arrival_time", &transmit_arrival_time, py::arg("transmit"), py::arg("point"),
        py::arg("speed_of_sound"));
  m.def("transmit_sensitivity", &transmit_sensitivity, py::arg("probe"), py::arg("transmit"),
        py::arg("point"), py::arg("wavelength"));
  m.def("element_directivity", &element_directivity, py::arg("element_width"),
        py::arg("wavelength"), py::arg("sin_theta"));
  m.def("dynamic_aperture_weight", &dynamic_aperture_weight, py::arg("element"), py::arg("point"),
        py::arg("f_number"), py::arg("window"), py::arg("tukey_alpha") = 0.5f);
  m.def("window_value", &window_value, py::arg("window"), py::arg("u"),
        py::arg("tukey_alpha") = 0.5f);

  py::class_<ScanGrid>(m, "ScanGrid")
      .def(py::init<>())
      .def_static("cartesian", &ScanGrid::cartesian, py::arg("x_min"), py::arg("x_max"),
                  py::arg("nx"), py::arg("z_min"), py::arg("z_max"), py::arg("nz"))
      .def_static("sector", &ScanGrid::sector, py::arg("theta_min"), py::arg("theta_max"),
                  py::arg("n_theta"), py::arg("r_min"), py::arg("r_max"), py::arg("n_r"),
                  py::arg("apex"))
      .def_readonly("type", &ScanGrid::type)
      .def_readonly("n_axis0", &ScanGrid::n_axis0)
      .def_readonly("n_axis1", &ScanGrid::n_axis1)
      .def_readonly("apex", &ScanGrid::apex)
      .def_property_readonly("axis0", [](ScanGrid& g) { return copy_vector(g.axis0, {g.n_axis0}); })
      .def_property_readonly("axis1", [](ScanGrid& g) { return copy_vector(g.axis1, {g.n_axis1}); })
      .def_property_readonly("shape",
                             [](const ScanGrid& g) { return py::make_tuple(g.n_axis0, g.n_axis1); })
      .def_property_readonly("n_pixels", &ScanGrid::n_pixels)
      .def_property_readonly("points", [](ScanGrid& g) {
        py::array_t<real> a({static_cast<py::ssize_t>(g.n_pixels()), py::ssize_t(3)});
        auto r = a.mutable_unchecked<2>();
        for (int k = 0; k < g.n_pixels(); ++k) {
          r(k, 0) = g.points[k].x;
          r(k, 1) = g.points[k].y;
          r(k, 2) = g.points[k].z;
        }
        return a;
      });

  py::class_<ChannelData>(m, "ChannelData")
      .def(py::init<>())
      .def_readwrite("probe", &ChannelData::probe)
      .def_readwrite("medium", &ChannelData::medium)
      .def_readwrite("acq", &ChannelData::acq)
      .def_readwrite("transmits", &ChannelData::transmits)
      .def_readonly("is_iq", &ChannelData::is_iq)
      .def_readonly("demodulation_frequency", &ChannelData::demodulation_frequency)
      .def_property_readonly("n_events", &ChannelData::n_events)
      .def_property_readonly("n_channels", &ChannelData::n_channels)
      .def_property_readonly("n_samples", &ChannelData::n_samples)
      .def_property_readonly("shape",
                             [](const ChannelData& d) {
                               return py::make_tuple(d.n_events(), d.n_channels(), d.n_samples());
                             })
      .def_property_readonly("rf",
                             [](py::object self) {
                               ChannelData& d = self.cast<ChannelData&>();
                               if (d.is_iq) throw std::runtime_error("ChannelData holds IQ, not RF");
                               return view<real>(d.rf.data(),
                                                 {d.n_events(), d.n_channels(), d.n_samples()}, self);
                             })
      .def_property_readonly("iq",
                             [](py::object self) {
                               ChannelData& d = self.cast<ChannelData&>();
                               if (!d.is_iq) throw std::runtime_error("ChannelData holds RF, not IQ");
                               return view<std::complex<real>>(
                                   d.iq.data(), {d.n_events(), d.n_channels(), d.n_samples()}, self);
                             })
      .def("validate", &ChannelData::validate)
      .def_static(
          "from_rf",
          [](py::array_t<real, py::array::c_style | py::array::forcecast> arr, Probe probe,
             Medium medium, Acquisition acq, std::vector<Transmit> transmits) {
            if (arr.ndim() != 3)
              throw std::invalid_argument(
                  "from_rf expects an array of shape (events, channels, samples)");
            ChannelData d;
            d.probe = probe;
            d.medium = medium;
            d.acq = acq;
            d.acq.n_samples = static_cast<int>(arr.shape(2));
            d.transmits = std::move(transmits);
            d.is_iq = false;
            d.allocate();
            if (static_cast<std::size_t>(arr.size()) != d.size())
              throw std::invalid_argument("from_rf: array shape does not match probe/transmits");
            std::memcpy(d.rf.data(), arr.data(), d.size() * sizeof(real));
            return d;
          },
          py::arg("array"), py::arg("probe"), py::arg("medium"), py::arg("acquisition"),
          py::arg("transmits"))
      .def_static(
          "from_iq",
          [](py::array_t<std::complex<real>, py::array::c_style | py::array::forcecast> arr,
             Probe probe, Medium medium, Acquisition acq, std::vector<Transmit> transmits,
             real demodulation_frequency) {
            if (arr.ndim() != 3)
              throw std::invalid_argument(
                  "from_iq expects an array of shape (events, channels, samples)");
            ChannelData d;
            d.probe = probe;
            d.medium = medium;
            d.acq = acq;
            d.acq.n_samples = static_cast<int>(arr.shape(2));
            d.transmits = std::move(transmits);
            d.is_iq = true;
            d.demodulation_frequency = demodulation_frequency;
            d.allocate();
            if (static_cast<std::size_t>(arr.size()) != d.size())
              throw std::invalid_argument("from_iq: array shape does not match probe/transmits");
            std::memcpy(static_cast<void*>(d.iq.data()), arr.data(), d.size() * sizeof(cplx));
            return d;
          },
          py::arg("array"), py::arg("probe"), py::arg("medium"), py::arg("acquisition"),
          py::arg("transmits"), py::arg("demodulation_frequency"));

  py::class_<Frame>(m, "Frame")
      .def(py::init<>())
      .def_readonly("grid", &Frame::grid)
      .def_readwrite("demodulation_frequency", &Frame::demodulation_frequency)
      .def_readwrite("timestamp", &Frame::timestamp)
      .def_readwrite("sequence", &Frame::sequence)
      .def_property_readonly("data",
                             [](py::object self) {
                               Frame& f = self.cast<Frame&>();
                               return view<std::complex<real>>(
                                   f.data.data(), {f.grid.n_axis0, f.grid.n_axis1}, self);
                             })
      .def_property_readonly("shape", [](const Frame& f) {
        return py::make_tuple(f.grid.n_axis0, f.grid.n_axis1);
      });

  py::class_<Scatterer>(m, "Scatterer")
      .def(py::init<>())
      .def_readwrite("position", &Scatterer::position)
      .def_readwrite("amplitude", &Scatterer::amplitude)
      .def_readwrite("velocity", &Scatterer::velocity);

  py::class_<Phantom>(m, "Phantom")
      .def(py::init<>())
      .def("add", &Phantom::add)
      .def("__len__", &Phantom::size)
      .def_static("point_targets", &Phantom::point_targets, py::arg("positions"),
                  py::arg("amplitude") = 1.0f)
      .def_static("speckle_box", &Phantom::speckle_box, py::arg("x_min"), py::arg("x_max"),
                  py::arg("z_min"), py::arg("z_max"), py::arg("y_half_thickness"),
                  py::arg("n_scatterers"), py::arg("seed") = 1234u,
                  py::arg("amplitude_sigma") = 1.0f)
      .def_static("recommended_scatterer_count", &Phantom::recommended_scatterer_count,
                  py::arg("probe"), py::arg("medium"), py::arg("x_span"), py::arg("z_span"),
                  py::arg("y_span"), py::arg("f_number") = 2.0f, py::arg("per_cell") = 10.0f)
      .def("scale_amplitude_in_circle", &Phantom::scale_amplitude_in_circle, py::arg("center"),
           py::arg("radius"), py::arg("scale"))
      .def("add_flow_tube", &Phantom::add_flow_tube, py::arg("center"), py::arg("radius"),
           py::arg("length"), py::arg("axis_angle"), py::arg("peak_velocity"),
           py::arg("n_scatterers"), py::arg("seed") = 99u, py::arg("amplitude") = 0.05f)
      .def("extend",
           [](Phantom& self, const Phantom& other) {
             self.scatterers.insert(self.scatterers.end(), other.scatterers.begin(),
                                    other.scatterers.end());
           })
      .def_property(
          "positions",
          [](const Phantom& p) {
            py::array_t<real> a({static_cast<py::ssize_t>(p.scatterers.size()), py::ssize_t(3)});
            auto r = a.mutable_unchecked<2>();
            for (std::size_t i = 0; i < p.scatterers.size(); ++i) {
              r(i, 0) = p.scatterers[i].position.x;
              r(i, 1) = p.scatterers[i].position.y;
              r(i, 2) = p.scatterers[i].position.z;
            }
            return a;
          },
          [](Phantom& p, py::array_t<real, py::array::c_style | py::array::forcecast> a) {
            if (a.ndim() != 2 || a.shape(1) != 3)
              throw std::invalid_argument("positions must have shape (n, 3)");
            p.scatterers.resize(static_cast<std::size_t>(a.shape(0)));
            auto r = a.unchecked<2>();
            for (py::ssize_t i = 0; i < a.shape(0); ++i)
              p.scatterers[i].position = Vec3(r(i, 0), r(i, 1), r(i, 2));
          })
      .def_property(
          "amplitudes",
          [](const Phantom& p) {
            py::array_t<real> a(static_cast<py::ssize_t>(p.scatterers.size()));
            auto r = a.mutable_unchecked<1>();
            for (std::size_t i = 0; i < p.scatterers.size(); ++i) r(i) = p.scatterers[i].amplitude;
            return a;
          },
          [](Phantom& p, py::array_t<real, py::array::c_style | py::array::forcecast> a) {
            if (static_cast<std::size_t>(a.size()) != p.scatterers.size())
              throw std::invalid_argument("amplitudes length must match the scatterer count");
            auto r = a.unchecked<1>();
            for (py::ssize_t i = 0; i < a.size(); ++i) p.scatterers[i].amplitude = r(i);
          })
      .def_property(
          "velocities",
          [](const Phantom& p) {
            py::array_t<real> a({static_cast<py::ssize_t>(p.scatterers.size()), py::ssize_t(3)});
            auto r = a.mutable_unchecked<2>();
            for (std::size_t i = 0; i < p.scatterers.size(); ++i) {
              r(i, 0) = p.scatterers[i].velocity.x;
              r(i, 1) = p.scatterers[i].velocity.y;
              r(i, 2) = p.scatterers[i].velocity.z;
            }
            return a;
          },
          [](Phantom& p, py::array_t<real, py::array::c_style | py::array::forcecast> a) {
            if (a.ndim() != 2 || a.shape(1) != 3 ||
                static_cast<std::size_t>(a.shape(0)) != p.scatterers.size())
              throw std::invalid_argument("velocities must have shape (n, 3) matching the phantom");
            auto r = a.unchecked<2>();
            for (py::ssize_t i = 0; i < a.shape(0); ++i)
              p.scatterers[i].velocity = Vec3(r(i, 0), r(i, 1), r(i, 2));
          })
      .def("copy", [](const Phantom& p) { return Phantom(p); });

  py::class_<SimulatorConfig>(m, "SimulatorConfig")
      .def(py::init<>())
      .def_readwrite("pulse_cycles", &SimulatorConfig::pulse_cycles)
      .def_readwrite("tx_window", &SimulatorConfig::tx_window)
      .def_readwrite("tx_f_number", &SimulatorConfig::tx_f_number)
      .def_readwrite("oversampling", &SimulatorConfig::oversampling)
      .def_readwrite("full_transmit_diffraction", &SimulatorConfig::full_transmit_diffraction)
      .def_readwrite("incident_trim_db", &SimulatorConfig::incident_trim_db)
      .def_readwrite("max_incident_window_us", &SimulatorConfig::max_incident_window_us)
      .def_readwrite("apply_attenuation", &SimulatorConfig::apply_attenuation)
      .def_readwrite("apply_directivity", &SimulatorConfig::apply_directivity)
      .def_readwrite("apply_spreading", &SimulatorConfig::apply_spreading)
      .def_readwrite("noise_db", &SimulatorConfig::noise_db)
      .def_readwrite("noise_seed", &SimulatorConfig::noise_seed)
      .def_readwrite("enable_motion", &SimulatorConfig::enable_motion);

  py::class_<Simulator>(m, "Simulator")
      .def(py::init<Probe, Medium, Acquisition, SimulatorConfig>(), py::arg("probe"),
           py::arg("medium"), py::arg("acquisition"), py::arg("config") = SimulatorConfig{})
      .def("make_plane_waves", &Simulator::make_plane_waves, py::arg("angles"))
      .def("make_focused_scan", &Simulator::make_focused_scan, py::arg("n_lines"),
           py::arg("x_min"), py::arg("x_max"), py::arg("focal_depth"), py::arg("angle") = 0.0f)
      .def("make_diverging_waves", &Simulator::make_diverging_waves, py::arg("angles"),
           py::arg("virtual_source_depth"))
      .def_static("repeat", &Simulator::repeat, py::arg("transmits"), py::arg("n_ensemble"))
      .def("simulate", &Simulator::simulate, py::arg("transmits"), py::arg("phantom"),
           py::call_guard<py::gil_scoped_release>())
      .def("simulate_gpu", &Simulator::simulate_gpu, py::arg("transmits"), py::arg("phantom"),
           py::call_guard<py::gil_scoped_release>())
      .def_property_readonly("probe", py::overload_cast<>(&Simulator::probe, py::const_))
      .def_property_readonly("medium", py::overload_cast<>(&Simulator::medium, py::const_))
      .def_property_readonly("acquisition",
                             py::overload_cast<>(&Simulator::acquisition, py::const_))
      .def("pulse", [](const Simulator& s) {
        std::vector<real> p = s.pulse(nullptr);
        return copy_vector(p, {static_cast<py::ssize_t>(p.size())});
      });

  py::class_<DemodConfig>(m, "DemodConfig")
      .def(py::init<>())
      .def_readwrite("demodulation_frequency", &DemodConfig::demodulation_frequency)
      .def_readwrite("cutoff_fraction", &DemodConfig::cutoff_fraction)
      .def_readwrite("filter_taps", &DemodConfig::filter_taps)
      .def_readwrite("decimation", &DemodConfig::decimation);

  m.def("demodulate", &demodulate, py::arg("channel_data"), py::arg("config") = DemodConfig{},
        py::call_guard<py::gil_scoped_release>());
  m.def(
      "design_lowpass",
      [](int n, real fc) {
        std::vector<real> h = design_lowpass(n, fc);
        return copy_vector(h, {static_cast<py::ssize_t>(h.size())});
      },
      py::arg("n_taps"), py::arg("normalized_cutoff"));

  py::class_<BeamformerConfig>(m, "BeamformerConfig")
      .def(py::init<>())
      .def_readwrite("f_number", &BeamformerConfig::f_number)
      .def_readwrite("rx_window", &BeamformerConfig::rx_window)
      .def_readwrite("tukey_alpha", &BeamformerConfig::tukey_alpha)
      .def_readwrite("combiner", &BeamformerConfig::combiner)
      .def_readwrite("compounding", &BeamformerConfig::compounding)
      .def_readwrite("transmit_weighting", &BeamformerConfig::transmit_weighting)
      .def_readwrite("coherence_factor", &BeamformerConfig::coherence_factor)
      .def_readwrite("mv_subarray", &BeamformerConfig::mv_subarray)
      .def_readwrite("mv_diagonal_loading", &BeamformerConfig::mv_diagonal_loading)
      .def_readwrite("mv_range_average", &BeamformerConfig::mv_range_average)
      .def_readwrite("mv_precision", &BeamformerConfig::mv_precision)
      .def_readwrite("mv_refine", &BeamformerConfig::mv_refine)
      .def_readwrite("tile_slots", &BeamformerConfig::tile_slots)
      .def_readwrite("tile_mv_occupancy", &BeamformerConfig::tile_mv_occupancy);

  py::class_<Beamformer>(m, "Beamformer")
      .def(py::init<BeamformerConfig>(), py::arg("config") = BeamformerConfig{})
      .def("beamform", &Beamformer::beamform, py::arg("channel_data"), py::arg("grid"),
           py::arg("event_begin") = 0, py::arg("event_end") = -1,
           py::call_guard<py::gil_scoped_release>())
      .def("beamform_sequence", &Beamformer::beamform_sequence, py::arg("channel_data"),
           py::arg("grid"), py::arg("events_per_frame"), py::call_guard<py::gil_scoped_release>())
      .def_property("config", py::overload_cast<>(&Beamformer::config, py::const_),
                    [](Beamformer& b, const BeamformerConfig& c) { b.config() = c; });

  py::class_<PipelineConfig>(m, "PipelineConfig")
      .def(py::init<>())
      .def_readwrite("demod", &PipelineConfig::demod)
      .def_readwrite("beamformer", &PipelineConfig::beamformer)
      .def_readwrite("dynamic_range_db", &PipelineConfig::dynamic_range_db)
      .def_readwrite("gain_db", &PipelineConfig::gain_db)
      .def_readwrite("apply_tgc", &PipelineConfig::apply_tgc)
      .def_readwrite("reference_smoothing", &PipelineConfig::reference_smoothing);

  py::class_<PipelineStats>(m, "PipelineStats")
      .def_readonly("demodulate_ms", &PipelineStats::demodulate_ms)
      .def_readonly("beamform_ms", &PipelineStats::beamform_ms)
      .def_readonly("postprocess_ms", &PipelineStats::postprocess_ms)
      .def_readonly("total_ms", &PipelineStats::total_ms)
      .def_property_readonly("fps", &PipelineStats::frames_per_second)
      .def("__repr__", [](const PipelineStats& s) {
        char buf[192];
        std::snprintf(buf, sizeof(buf),
                      "PipelineStats(demod=%.1fms, beamform=%.1fms, post=%.1fms, total=%.1fms, "
                      "%.1f fps)",
                      s.demodulate_ms, s.beamform_ms, s.postprocess_ms, s.total_ms,
                      s.frames_per_second());
        return std::string(buf);
      });

  py::class_<ImagingPipeline>(m, "ImagingPipeline")
      .def(py::init<ScanGrid, PipelineConfig>(), py::arg("grid"),
           py::arg("config") = PipelineConfig{})
      .def("process_frame", &ImagingPipeline::process_frame, py::arg("channel_data"),
           py::call_guard<py::gil_scoped_release>())
      .def(
          "process_display",
          [](ImagingPipeline& p, const ChannelData& d) {
            std::vector<real> img;
            {
              py::gil_scoped_release release;
              img = p.process_display(d);
            }
            return copy_vector(img, {p.grid().n_axis0, p.grid().n_axis1});
          },
          py::arg("channel_data"))
      .def(
          "to_display",
          [](ImagingPipeline& p, const Frame& f, const Medium& med) {
            std::vector<real> img;
            {
              py::gil_scoped_release release;
              img = p.to_display(f, med);
            }
            return copy_vector(img, {p.grid().n_axis0, p.grid().n_axis1});
          },
          py::arg("frame"), py::arg("medium") = Medium{})
      .def("reset", &ImagingPipeline::reset)
      .def(
          "set_apodization",
          [](ImagingPipeline& p, py::array_t<float, py::array::c_style | py::array::forcecast> a) {
            p.set_apodization(std::vector<float>(a.data(), a.data() + a.size()));
          },
          py::arg("params"))
      .def_property_readonly("has_apodization", &ImagingPipeline::has_apodization)
      .def_property_readonly("grid", &ImagingPipeline::grid)
      .def_property_readonly("stats", &ImagingPipeline::stats)
      .def_property_readonly("frame_count", &ImagingPipeline::frame_count)
      .def_property("config", py::overload_cast<>(&ImagingPipeline::config, py::const_),
                    [](ImagingPipeline& p, const PipelineConfig& c) { p.config() = c; });

  py::class_<DeviceBuffer>(m, "DeviceBuffer")
      .def(py::init<std::size_t>(), py::arg("nbytes"))
      .def_property_readonly("ptr", &DeviceBuffer::ptr)
      .def_property_readonly("nbytes", &DeviceBuffer::nbytes)
      .def("upload", &DeviceBuffer::upload, py::arg("array"))
      .def("download", &DeviceBuffer::download, py::arg("dtype"), py::arg("shape"))
      .def("zero", &DeviceBuffer::zero);

  py::class_<DeviceRunner>(m, "DeviceRunner")
      .def(py::init<const ScanGrid&, const PipelineConfig&, const ChannelData&>(), py::arg("grid"),
           py::arg("config"), py::arg("like"))
      .def("run", &DeviceRunner::run, py::arg("input_ptr"), py::arg("frame_ptr"),
           py::call_guard<py::gil_scoped_release>())
      .def("run_noalloc", &DeviceRunner::run_noalloc, py::arg("input_ptr"), py::arg("frame_ptr"),
           py::call_guard<py::gil_scoped_release>())
      .def("synchronize", &DeviceRunner::synchronize, py::call_guard<py::gil_scoped_release>())
      .def("capture", &DeviceRunner::capture, py::arg("input_ptr"), py::arg("frame_ptr"),
           py::call_guard<py::gil_scoped_release>())
      .def("run_graph", &DeviceRunner::run_graph, py::call_guard<py::gil_scoped_release>())
      .def("capture_host", &DeviceRunner::capture_host, py::arg("channel_data"))
      .def("graph_frame", &DeviceRunner::graph_frame)
      .def_property_readonly("stream", &DeviceRunner::stream)
      .def("record", &DeviceRunner::record, py::arg("slot"))
      .def("elapsed", &DeviceRunner::elapsed, py::arg("a"), py::arg("b"))
      .def("profile", &DeviceRunner::profile, py::arg("on"))
      .def("profile_report", &DeviceRunner::profile_report)
      .def_property_readonly("input_shape", &DeviceRunner::input_shape)
      .def_property_readonly("is_rf", &DeviceRunner::is_rf)
      .def_property_readonly("iq_samples", &DeviceRunner::iq_samples)
      .def_property_readonly("frame_pixels", &DeviceRunner::frame_pixels)
      .def("set_apodization", &DeviceRunner::set_apodization, py::arg("params"))
      .def("get_apodization", &DeviceRunner::get_apodization)
      .def("mv_weights", &DeviceRunner::mv_weights, py::arg("channel_data"))
      .def("dataset_reserve", &DeviceRunner::dataset_reserve, py::arg("capacity"))
      .def_property_readonly("dataset_size", &DeviceRunner::dataset_size)
      .def("dataset_append", &DeviceRunner::dataset_append, py::arg("channel_data"),
           py::arg("pixels"))
      .def("train", &DeviceRunner::train, py::arg("n_steps"), py::arg("batch") = 4096,
           py::arg("lr") = 1e-3f, py::arg("beta1") = 0.9f, py::arg("beta2") = 0.999f,
           py::arg("eps") = 1e-8f, py::arg("weight_decay") = 0.0f, py::arg("seed") = 0ull)
      .def("gradient", &DeviceRunner::gradient, py::arg("x"), py::arg("g"), py::arg("mask"),
           py::arg("sw"))
      .def("autotune", &DeviceRunner::autotune, py::arg("input_ptr"), py::arg("frame_ptr"),
           py::arg("reps") = 3)
      .def("time_host", &DeviceRunner::time_host, py::arg("channel_data"), py::arg("reps") = 10,
           py::call_guard<py::gil_scoped_release>())
      .def("autotune_host", &DeviceRunner::autotune_host, py::arg("channel_data"),
           py::arg("reps") = 3);

  py::class_<FrameSequence>(m, "FrameSequence")
      .def(py::init<const std::vector<ChannelData*>&, const std::vector<ScanGrid*>&,
                    const std::vector<PipelineConfig>&>(),
           py::arg("channel_data"), py::arg("grids"), py::arg("configs"))
      .def("clear", &FrameSequence::clear)
      .def("frames", &FrameSequence::frames)
      .def("tiles", &FrameSequence::tiles)
      .def("launch", &FrameSequence::launch, py::arg("graph") = true,
           py::call_guard<py::gil_scoped_release>())
      .def("synchronize", &FrameSequence::synchronize, py::call_guard<py::gil_scoped_release>())
      .def("record", &FrameSequence::record, py::arg("slot"))
      .def("elapsed", &FrameSequence::elapsed, py::arg("a"), py::arg("b"))
      .def("__len__", &FrameSequence::size);
  // --- doppler / elastography (reference :486-545) -------------------------
  py::class_<DopplerConfig>(m, "DopplerConfig")
      .def(py::init<>())
      .def_readwrite("wall_filter", &DopplerConfig::wall_filter)
      .def_readwrite("polynomial_order", &DopplerConfig::polynomial_order)
      .def_readwrite("loupas", &DopplerConfig::loupas)
      .def_readwrite("kernel_axial", &DopplerConfig::kernel_axial)
      .def_readwrite("kernel_lateral", &DopplerConfig::kernel_lateral)
      .def_readwrite("power_threshold_db", &DopplerConfig::power_threshold_db)
      .def_readwrite("coherence_threshold", &DopplerConfig::coherence_threshold);

  py::class_<ColorFlowMap>(m, "ColorFlowMap")
      .def_readonly("grid", &ColorFlowMap::grid)
      .def_readonly("nyquist_velocity", &ColorFlowMap::nyquist_velocity)
      .def_property_readonly("velocity",
                             [](py::object self) {
                               ColorFlowMap& c = self.cast<ColorFlowMap&>();
                               return view<real>(c.velocity.data(), {c.grid.n_axis0, c.grid.n_axis1}, self);
                             })
      .def_property_readonly("power",
                             [](py::object self) {
                               ColorFlowMap& c = self.cast<ColorFlowMap&>();
                               return view<real>(c.power.data(), {c.grid.n_axis0, c.grid.n_axis1}, self);
                             })
      .def_property_readonly("variance",
                             [](py::object self) {
                               ColorFlowMap& c = self.cast<ColorFlowMap&>();
                               return view<real>(c.variance.data(), {c.grid.n_axis0, c.grid.n_axis1}, self);
                             })
      .def_property_readonly("valid", [](py::object self) {
        ColorFlowMap& c = self.cast<ColorFlowMap&>();
        return view<std::uint8_t>(c.valid.data(), {c.grid.n_axis0, c.grid.n_axis1}, self);
      });

  m.def("color_flow", &color_flow, py::arg("ensemble"), py::arg("prf"), py::arg("speed_of_sound"),
        py::arg("config") = DopplerConfig{}, py::call_guard<py::gil_scoped_release>());
  m.def("apply_wall_filter", &apply_wall_filter, py::arg("ensemble"), py::arg("type"),
        py::arg("polynomial_order") = 1, py::call_guard<py::gil_scoped_release>());
  m.def(
      "power_doppler",
      [](const std::vector<Frame>& ens, WallFilter type, int order) {
        std::vector<real> p;
        {
          py::gil_scoped_release release;
          p = power_doppler(ens, type, order);
        }
        if (ens.empty()) return copy_vector(p, {0});
        return copy_vector(p, {ens.front().grid.n_axis0, ens.front().grid.n_axis1});
      },
      py::arg("ensemble"), py::arg("type") = WallFilter::PolynomialRegression,
      py::arg("polynomial_order") = 1);
  m.def(
      "polynomial_reject_matrix",
      [](int n, int order) { return copy_vector(polynomial_reject_matrix(n, order), {n, n}); },
      py::arg("n"), py::arg("order"));

  py::class_<DisplacementMap>(m, "DisplacementMap")
      .def_readonly("grid", &DisplacementMap::grid)
      .def_property_readonly("axial",
                             [](py::object self) {
                               DisplacementMap& d = self.cast<DisplacementMap&>();
                               return view<real>(d.axial.data(), {d.grid.n_axis0, d.grid.n_axis1}, self);
                             })
      .def_property_readonly("correlation", [](py::object self) {
        DisplacementMap& d = self.cast<DisplacementMap&>();
        return view<real>(d.correlation.data(), {d.grid.n_axis0, d.grid.n_axis1}, self);
      });

  m.def("estimate_axial_displacement", &estimate_axial_displacement, py::arg("before"),
        py::arg("after"), py::arg("speed_of_sound"), py::arg("kernel_axial") = 8,
        py::arg("kernel_lateral") = 2, py::arg("max_lag") = 8,
        py::call_guard<py::gil_scoped_release>());
  m.def(
      "axial_strain",
      [](const DisplacementMap& d, int fit_length) {
        std::vector<real> st;
        {
          py::gil_scoped_release release;
          st = axial_strain(d, fit_length);
        }
        return copy_vector(st, {d.grid.n_axis0, d.grid.n_axis1});
      },
      py::arg("displacement"), py::arg("fit_length") = 15);

  py::class_<DopplerRunner>(m, "DopplerRunner")
      .def(py::init<const ScanGrid&, int, real, real, real, const DopplerConfig&>(),
           py::arg("grid"), py::arg("n_frames"), py::arg("prf"), py::arg("speed_of_sound"),
           py::arg("demodulation_frequency"), py::arg("config") = DopplerConfig{})
      .def("color_flow", &DopplerRunner::color_flow, py::arg("ensemble_ptr"),
           py::arg("velocity_ptr"), py::arg("power_ptr"), py::arg("variance_ptr"),
           py::arg("valid_ptr"))
      .def("power_doppler", &DopplerRunner::power_doppler, py::arg("ensemble_ptr"),
           py::arg("power_ptr"))
      .def("wall_filter", &DopplerRunner::wall_filter, py::arg("ensemble_ptr"), py::arg("out_ptr"))
      .def("synchronize", &DopplerRunner::synchronize, py::call_guard<py::gil_scoped_release>())
      .def("record", &DopplerRunner::record, py::arg("slot"))
      .def("elapsed", &DopplerRunner::elapsed, py::arg("a"), py::arg("b"))
      .def("profile", &DopplerRunner::profile, py::arg("on"))
      .def("profile_report", &DopplerRunner::profile_report)
      .def_property_readonly("n_frames", &DopplerRunner::n_frames);

  m.attr("APOD_PARAMS") = USX_APOD_PARAMS;

  m.def("build_info", []() { return std::string(usx_build_info()); });
  m.attr("__version__") = "0.1.0";
}
