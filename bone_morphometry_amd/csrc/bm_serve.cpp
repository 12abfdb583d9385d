// bm_serve.cpp -- batched (hipGraph) and serving forms of analyze.
#include "bm_serve.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "bm_analyze.h"

namespace bm {

static size_t align256(size_t v) { return (v + 255) / 256 * 256; }

namespace {
// analyze's checks of one volume, as capture_errors callbacks.
struct CheckCtx {
  uint64_t w, h, d;
  const double* spacing;
  int mode;
  Dims dims;
};
void check_volume(void* p) {
  CheckCtx* c = static_cast<CheckCtx*>(p);
  check_spacing(c->spacing);
  check_mode(c->mode);
  c->dims = device_dims(c->w, c->h, c->d);
}
struct FinishCtx {
  const ItemResult* bone;
  const ItemResult* marrow;
  uint64_t n;
  const double* spacing;
  int mode;
  bm_thickness_spacing* out;
  const StatsResult* seq;
};
void finish_volume(void* p) {
  FinishCtx* c = static_cast<FinishCtx*>(p);
  analyze_finish(*c->bone, *c->marrow, c->n, c->spacing, c->mode, c->out, c->seq);
}

VolumeResult finish(const ItemResult& bone, const ItemResult& marrow, const Dims& dims,
                    const double spacing[3], int mode, const StatsResult* seq = nullptr) {
  VolumeResult r;
  FinishCtx c{&bone, &marrow, dims.n(), spacing, mode, &r.metrics, seq};
  r.status = capture_errors(finish_volume, &c, &r.message);
  if (r.status != BM_OK) r.metrics = bm_thickness_spacing{};
  return r;
}
}  // namespace

// ---------------------------------------------------------------------------
Batch::Batch(const bm_batch_item* items, size_t count, bool use_graph)
    : items_(items, items + count), first_item_(count, -1), early_(count),
      seq_(stats_sequential()) {
  std::vector<ItemSpec> specs;
  for (size_t i = 0; i < count; ++i) {
    const bm_batch_item& it = items_[i];
    CheckCtx c{it.width, it.height, it.depth, it.voxel_spacing, it.feature_mode, Dims()};
    early_[i].status = capture_errors(check_volume, &c, &early_[i].message);
    if (early_[i].status != BM_OK) continue;
    if (!it.d_bone || !it.d_roi) {
      early_[i].status = BM_ERR_INVALID;
      early_[i].message = "null volume";
      continue;
    }
    first_item_[i] = (int)specs.size();
    analyze_items(it.d_bone, it.d_roi, c.dims, it.voxel_spacing, it.feature_mode, specs, seq_);
  }
  BM_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  if (specs.empty()) return;
  plan_.build(specs, stream_);
  if (!use_graph) return;
  // Capture the launch sequence once (no profiling events inside the graph).
  profile_suspend(true);
  hipError_t e = hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    try {
      plan_.launch(stream_);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream_, &g);
      if (g) (void)hipGraphDestroy(g);
      profile_suspend(false);
      throw;
    }
    e = hipStreamEndCapture(stream_, &graph_);
  }
  profile_suspend(false);
  BM_HIP(e);
  BM_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
}

Batch::~Batch() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Batch::run(std::vector<VolumeResult>& out) {
  if (plan_.size()) {
    if (exec_) BM_HIP(hipGraphLaunch(exec_, stream_));
    else plan_.launch(stream_);
    plan_.fetch(stream_);
  }
  out.assign(items_.size(), VolumeResult());
  for (size_t i = 0; i < items_.size(); ++i) {
    if (first_item_[i] < 0) {
      out[i] = early_[i];
      continue;
    }
    const bm_batch_item& it = items_[i];
    Dims dims;
    dims.w = (uint32_t)it.width;
    dims.h = (uint32_t)it.height;
    dims.d = (uint32_t)it.depth;
    StatsResult sums[2];
    const bool replayed =
        seq_ && replay_sums(plan_, (size_t)first_item_[i], (float)it.voxel_spacing[0], stream_, sums);
    out[i] = finish(plan_.result(first_item_[i]), plan_.result(first_item_[i] + 1), dims,
                    it.voxel_spacing, it.feature_mode, replayed ? sums : nullptr);
  }
}

// ---------------------------------------------------------------------------
Server::Server(uint64_t budget_bytes, int max_batch, int slots) {
  device_ = current_device();
  t0_ = std::chrono::steady_clock::now();
  budget_ = budget_bytes;
  max_batch_ = std::max(1, max_batch);
  const int n = std::max(1, slots);
  size_t free_b = 0, total_b = 0;
  BM_HIP(hipMemGetInfo(&free_b, &total_b));
  slot_cap_ = std::min<uint64_t>(budget_ / n, (uint64_t)(free_b / 10 * 9) / n);
  if (const char* env = std::getenv("BM_SERVER_FAIL_BATCH")) fail_batch_ = std::strtol(env, nullptr, 10);
  for (int i = 0; i < n; ++i) {
    auto s = std::make_unique<Slot>();
    BM_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    BM_HIP(hipEventCreateWithFlags(&s->done, hipEventDisableTiming));
    s->plan.reserve_items((size_t)max_batch_ * 2);
    slots_.push_back(std::move(s));
  }
  thread_ = std::thread(&Server::worker, this);
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  if (thread_.joinable()) thread_.join();
  for (auto& s : slots_) {
    if (s->busy) (void)hipStreamSynchronize(s->stream);
    (void)hipEventDestroy(s->done);
    (void)hipStreamDestroy(s->stream);
  }
}

double Server::now() const {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

uint64_t Server::submit(const uint8_t* bone, const uint8_t* roi, uint64_t w, uint64_t h,
                        uint64_t d, const double spacing[3], int mode) {
  auto r = std::make_unique<Request>();
  r->bone = bone;
  r->roi = roi;
  std::memcpy(r->spacing, spacing, sizeof r->spacing);
  r->mode = mode;
  r->t_submit = now();
  CheckCtx c{w, h, d, spacing, mode, Dims()};
  r->result.status = capture_errors(check_volume, &c, &r->result.message);
  if (r->result.status == BM_OK && (!bone || !roi)) {
    r->result.status = BM_ERR_INVALID;
    r->result.message = "null volume";
  }
  if (r->result.status == BM_OK) {
    r->dims = c.dims;
    std::vector<ItemSpec> specs;
    analyze_items(nullptr, nullptr, r->dims, spacing, mode, specs);
    r->bytes = Plan::bytes_for(specs) + 2 * align256(r->dims.n());
    if (r->bytes > slot_cap_) {
      char buf[200];
      snprintf(buf, sizeof buf,
               "request needs %llu device bytes, more than a slot's share of the budget (%llu)",
               (unsigned long long)r->bytes, (unsigned long long)slot_cap_);
      r->result.status = BM_ERR_UNSUPPORTED;
      r->result.message = buf;
    }
  }
  std::lock_guard<std::mutex> g(mu_);
  if (!worker_error_.empty() && r->result.status == BM_OK) {
    r->result.status = BM_ERR_DEVICE;
    r->result.message = worker_error_;
  }
  const uint64_t t = next_ticket_++;
  r->ticket = t;
  Request* raw = r.get();
  if (r->result.status != BM_OK) {
    r->done = true;
    r->t_start = r->t_done = r->t_submit;
  }
  reqs_[t] = std::move(r);
  if (!raw->done) {
    queue_.push_back(raw);
    cv_work_.notify_one();
  } else {
    cv_done_.notify_all();
  }
  return t;
}

bool Server::cancel(uint64_t ticket) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = reqs_.find(ticket);
  if (it == reqs_.end()) throw Error("unknown ticket");
  Request* r = it->second.get();
  auto q = std::find(queue_.begin(), queue_.end(), r);
  if (r->done || q == queue_.end()) return false;  // launched or finished
  queue_.erase(q);
  r->result.status = BM_ERR_INVALID;
  r->result.message = "request cancelled";
  r->t_start = r->t_done = now();
  r->done = true;
  cv_done_.notify_all();
  return true;
}

bool Server::poll(uint64_t ticket, bool wait, VolumeResult* out, bm_request_timing* timing) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = reqs_.find(ticket);
  if (it == reqs_.end()) throw Error("unknown ticket");
  Request* r = it->second.get();
  if (!r->done && wait) cv_done_.wait(lk, [&] { return r->done; });
  if (!r->done) return false;
  if (out) *out = r->result;
  if (timing) {
    timing->submit_s = r->t_submit;
    timing->start_s = r->t_start;
    timing->done_s = r->t_done;
    timing->batch_size = r->batch_size;
    timing->pad = 0;
  }
  reqs_.erase(it);
  return true;
}

void Server::stats(bm_server_stats* out) {
  std::lock_guard<std::mutex> g(mu_);
  out->batches = n_batches_;
  out->requests_done = n_done_;
  out->max_batch = max_batch_seen_;
  out->sum_batch = sum_batch_;
  out->peak_device_bytes = peak_bytes_;
  out->max_inflight = max_inflight_;
  out->slots = (uint32_t)slots_.size();
  double busy = busy_total_;
  if (busy_since_ >= 0) busy += now() - busy_since_;
  out->busy_s = busy;
}

// Takes the next requests of the queue (first come, first served) that fit
// the slot and launches them; called with the lock held, releases it while
// uploading and launching.
bool Server::launch_slot(Slot& s, std::unique_lock<std::mutex>& lk) {
  s.reqs.clear();
  uint64_t bytes = 0;
  while (!queue_.empty() && (int)s.reqs.size() < max_batch_) {
    Request* r = queue_.front();
    if (bytes + r->bytes > slot_cap_) break;
    bytes += r->bytes;
    s.reqs.push_back(r);
    queue_.pop_front();
  }
  if (s.reqs.empty()) return false;
  s.busy = true;
  uint32_t inflight = 0;
  uint64_t inflight_bytes = 0;
  for (auto& o : slots_)
    if (o->busy) {
      ++inflight;
      for (Request* q : o->reqs) inflight_bytes += q->bytes;
    }
  if (inflight == 1) busy_since_ = now();
  max_inflight_ = std::max(max_inflight_, inflight);
  peak_bytes_ = std::max(peak_bytes_, inflight_bytes);
  lk.unlock();
  std::string err;
  int code = BM_OK;
  try {
    if (fail_batch_ >= 0 && launched_++ == fail_batch_)
      throw DeviceError("injected device allocation failure (BM_SERVER_FAIL_BATCH)");
    s.mem.reserve(slot_cap_);
    char* base = static_cast<char*>(s.mem.base());
    size_t off = 0;
    std::vector<ItemSpec> specs;
    s.first_item.clear();
    for (Request* r : s.reqs) {
      const size_t n = r->dims.n();
      uint8_t* d_bone = reinterpret_cast<uint8_t*>(base + off);
      uint8_t* d_roi = reinterpret_cast<uint8_t*>(base + off + align256(n));
      off += 2 * align256(n);
      BM_HIP(hipMemcpyAsync(d_bone, r->bone, n, hipMemcpyHostToDevice, s.stream));
      BM_HIP(hipMemcpyAsync(d_roi, r->roi, n, hipMemcpyHostToDevice, s.stream));
      s.first_item.push_back((int)specs.size());
      analyze_items(d_bone, d_roi, r->dims, r->spacing, r->mode, specs);
    }
    s.plan.build(specs, s.stream, base + off, slot_cap_ - off);
    const double t = now();
    for (Request* r : s.reqs) r->t_start = t;
    s.plan.launch(s.stream);
    BM_HIP(hipEventRecord(s.done, s.stream));
  } catch (const DeviceError& e) {
    code = BM_ERR_DEVICE;
    err = e.what();
  } catch (const std::exception& e) {
    code = BM_ERR_DEVICE;
    err = e.what();
  }
  lk.lock();
  ++n_batches_;
  sum_batch_ += s.reqs.size();
  max_batch_seen_ = std::max<uint64_t>(max_batch_seen_, s.reqs.size());
  if (code != BM_OK) {  // the batch never ran: its requests fail
    const double t = now();
    for (Request* r : s.reqs) {
      r->result.status = code;
      r->result.message = err;
      r->t_start = r->t_done = t;
      r->batch_size = (uint32_t)s.reqs.size();
      r->done = true;
      ++n_done_;
    }
    s.reqs.clear();
    s.busy = false;
    bool any = false;
    for (auto& o : slots_) any |= o->busy;
    if (!any && busy_since_ >= 0) {
      busy_total_ += now() - busy_since_;
      busy_since_ = -1;
    }
    cv_done_.notify_all();
  }
  return true;
}

void Server::finish_slot(Slot& s) {
  std::vector<VolumeResult> res(s.reqs.size());
  try {
    s.plan.fetch(s.stream);
    for (size_t k = 0; k < s.reqs.size(); ++k) {
      const Request* r = s.reqs[k];
      res[k] = finish(s.plan.result(s.first_item[k]), s.plan.result(s.first_item[k] + 1), r->dims,
                      r->spacing, r->mode);
    }
  } catch (const std::exception& e) {
    for (auto& v : res) {
      v.status = BM_ERR_DEVICE;
      v.message = e.what();
    }
  }
  const double t = now();
  std::lock_guard<std::mutex> g(mu_);
  for (size_t k = 0; k < s.reqs.size(); ++k) {
    Request* r = s.reqs[k];
    r->result = res[k];
    r->t_done = t;
    r->batch_size = (uint32_t)s.reqs.size();
    r->done = true;
    ++n_done_;
  }
  s.reqs.clear();
  s.busy = false;
  bool any = false;
  for (auto& o : slots_) any |= o->busy;
  if (!any && busy_since_ >= 0) {
    busy_total_ += t - busy_since_;
    busy_since_ = -1;
  }
  cv_done_.notify_all();
}

void Server::worker() {
  if (hipSetDevice(device_) != hipSuccess) {
    std::lock_guard<std::mutex> g(mu_);
    worker_error_ = "server worker cannot select its device";
    return;
  }
  std::unique_lock<std::mutex> lk(mu_);
  while (true) {
    bool progressed = false;
    for (auto& sp : slots_) {
      Slot& s = *sp;
      if (!s.busy) continue;
      const hipError_t q = hipEventQuery(s.done);
      if (q == hipErrorNotReady) continue;
      lk.unlock();
      finish_slot(s);  // also reports a failed batch (the fetch throws)
      lk.lock();
      progressed = true;
    }
    bool any_busy = false;
    for (auto& sp : slots_) any_busy |= sp->busy;
    if (stop_ && !any_busy) break;
    if (!stop_) {
      for (auto& sp : slots_) {
        if (sp->busy || queue_.empty()) continue;
        if (launch_slot(*sp, lk)) progressed = true;
      }
    }
    if (!progressed) {
      if (any_busy || stop_)
        cv_work_.wait_for(lk, std::chrono::microseconds(50));
      else
        cv_work_.wait(lk, [&] { return stop_ || !queue_.empty(); });
    }
  }
  // requests still queued at shutdown never ran
  for (Request* r : queue_) {
    r->result.status = BM_ERR_DEVICE;
    r->result.message = "server stopped";
    r->done = true;
  }
  queue_.clear();
  cv_done_.notify_all();
}

}  // namespace bm
