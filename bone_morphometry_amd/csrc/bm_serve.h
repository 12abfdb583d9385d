// bm_serve.h -- the batched and serving forms of analyze (thickness/spacing).
//
// Batch: many device-resident volumes of any shapes, each analyzed as by
// bm_analyze_thickness_spacing_device, as one plan: every stage is one launch
// over all the volumes' bone and marrow items, and the whole sequence is
// captured once into a hipGraph that each run() replays.
//
// Server: a request loop.  submit() queues a host-resident volume; a worker
// thread admits queued requests into a batch as soon as one of its slots is
// free and the batch fits the slot's share of the HBM budget, uploads them,
// and launches the batch on the slot's stream, so batches on different slots
// run concurrently on the device while new requests queue.  Results and
// per-request timestamps are collected with poll().
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bm.h"
#include "bm_engine.h"

namespace bm {

struct VolumeResult {
  bm_thickness_spacing metrics{};
  int status = BM_OK;
  std::string message;
};

class Batch {
 public:
  Batch(const bm_batch_item* items, size_t count, bool use_graph);
  ~Batch();
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;
  // Runs every volume (graph replay or direct launches) and waits.
  void run(std::vector<VolumeResult>& out);
  size_t device_bytes() const { return plan_.device_bytes(); }
  bool graph() const { return exec_ != nullptr; }

 private:
  std::vector<bm_batch_item> items_;
  std::vector<int> first_item_;       // plan item of each volume's bone phase, or -1
  std::vector<VolumeResult> early_;   // volumes rejected before the launch
  bool seq_ = false;                  // BM_STATS_SEQUENTIAL at creation
  Plan plan_;
  hipStream_t stream_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
};

class Server {
 public:
  Server(uint64_t budget_bytes, int max_batch, int slots);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;
  uint64_t submit(const uint8_t* bone, const uint8_t* roi, uint64_t w, uint64_t h, uint64_t d,
                  const double spacing[3], int mode);
  // Returns true when the request is done (its record is then released).
  bool poll(uint64_t ticket, bool wait, VolumeResult* out, bm_request_timing* timing);
  // Withdraws a request that is still queued (true); a request already
  // launched or done is left alone (false).
  bool cancel(uint64_t ticket);
  void stats(bm_server_stats* out);

 private:
  struct Request {
    uint64_t ticket;
    const uint8_t* bone;
    const uint8_t* roi;
    Dims dims;
    double spacing[3];
    int mode;
    size_t bytes;  // device bytes: inputs and plan buffers
    double t_submit = 0, t_start = 0, t_done = 0;
    bool done = false;
    uint32_t batch_size = 0;
    VolumeResult result;
  };
  struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    Arena mem;
    Plan plan;
    std::vector<Request*> reqs;
    std::vector<int> first_item;
    bool busy = false;
  };
  void worker();
  bool launch_slot(Slot& s, std::unique_lock<std::mutex>& lk);
  void finish_slot(Slot& s);
  double now() const;

  int device_ = 0;
  uint64_t budget_ = 0, slot_cap_ = 0;
  int max_batch_ = 0;
  std::vector<std::unique_ptr<Slot>> slots_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Request*> queue_;
  std::map<uint64_t, std::unique_ptr<Request>> reqs_;
  uint64_t next_ticket_ = 1;
  bool stop_ = false;
  // Test hook (BM_SERVER_FAIL_BATCH=k): the k-th batch launched (from 0)
  // fails as a device allocation failure would, before anything is uploaded.
  long fail_batch_ = -1;
  long launched_ = 0;
  std::thread thread_;
  std::string worker_error_;
  // statistics
  uint64_t n_batches_ = 0, n_done_ = 0, max_batch_seen_ = 0, sum_batch_ = 0;
  uint64_t peak_bytes_ = 0;
  uint32_t max_inflight_ = 0;
  double busy_since_ = -1, busy_total_ = 0;
  std::chrono::steady_clock::time_point t0_;
};

}  // namespace bm
