// Serving loop: DBA jobs (z-norm -> all-vs-all -> medoid -> DBA rounds to
// convergence, the performDBA pipeline for one cluster) submitted at any
// time, admitted in arrival order as soon as a slot and enough device
// memory are free, and run concurrently.
//
// Each slot is a worker thread with its own odba_ctx, hence its own HIP
// stream: the kernels of admitted jobs overlap on the device. A job that
// cannot fill the GPU (a few hundred pairs, or the one-wave-per-sequence
// DBA kernels) leaves compute units idle that the other jobs' persistent
// launches take up. Admission compares an estimate of the job's device
// working set (resident sequences, distance matrix, edge scratch, DBA
// traceback state and its path-word budget) with what is left of the
// server's HBM budget; each slot's context plans its DBA batches within
// budget / slots. A job's result does not depend on what runs beside it:
// every kernel is deterministic and the path mode is fixed (full matrix).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/odba/odba.h"

namespace {

size_t elem_size(int dtype) {
  switch (dtype) {
    case ODBA_I32: case ODBA_U32: case ODBA_F32: return 4;
    case ODBA_U64: case ODBA_F64: return 8;
  }
  return 0;
}

long long padded(long long L) { return ((L + 63) / 64) * 64 + 1024; }

struct Job {
  uint64_t id = 0;
  int dtype = 0;
  std::vector<std::vector<unsigned char>> data;
  std::vector<uint64_t> lens;
  int os = 0, oe = 0, norm = 1, max_rounds = 250;
  uint64_t footprint = 0;
  int state = ODBA_JOB_QUEUED;
  int rc = 0;
  std::string err;
  double t_submit = 0, t_admit = 0, t_done = 0;
  std::vector<unsigned char> centroid;
  uint64_t clen = 0;
  int medoid = -1, rounds = 0;
};

}  // namespace

struct odba_server {
  int device = 0;
  uint64_t budget = 0;
  int slots = 1;
  int cus = 256;
  std::vector<odba_ctx*> ctxs;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_admit, cv_done;
  std::deque<Job*> queue;
  std::map<uint64_t, std::unique_ptr<Job>> jobs;
  uint64_t in_use = 0;
  int running = 0;
  bool stop = false;
  uint64_t next_id = 1;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  std::string err;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  // Device bytes one job plans to hold while it runs.
  uint64_t footprint(const Job& j) const {
    const size_t es = elem_size(j.dtype);
    const long long n = (long long)j.lens.size();
    long long seq = 0, maxlen = 0, sumlen = 0;
    for (uint64_t L : j.lens) {
      seq += padded((long long)L);
      maxlen = std::max<long long>(maxlen, (long long)L);
      sumlen += (long long)L;
    }
    const long long npairs = n * (n - 1) / 2;
    const long long waves = std::min<long long>((npairs + 3) / 4, (long long)cus * 8) * 4;
    const uint64_t edge = (uint64_t)waves * 2 * (maxlen + 256) * es;
    const uint64_t dists = (uint64_t)npairs * es;
    const uint64_t runs = (uint64_t)n * maxlen * 16 + (uint64_t)padded(maxlen) * es * 4;
    const uint64_t words = (uint64_t)sumlen * (uint64_t)maxlen / 4 + (uint64_t)n * 64 * 4 * (maxlen / 1024 + 1);
    const uint64_t word_cap = budget / (uint64_t)slots / 4;
    return (uint64_t)seq * es * 2 + dists + edge + runs + std::min(words, word_cap);
  }

  bool can_admit() const {
    if (queue.empty() || running >= slots) return false;
    const Job* j = queue.front();
    return running == 0 || in_use + j->footprint <= budget;
  }

  void worker(int slot) {
    odba_ctx* c = ctxs[slot];
    for (;;) {
      Job* j = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_admit.wait(lk, [&] { return stop || can_admit(); });
        if (stop) return;
        j = queue.front();
        queue.pop_front();
        in_use += j->footprint;
        running++;
        j->state = ODBA_JOB_RUNNING;
        j->t_admit = now();
      }
      const int n = (int)j->lens.size();
      std::vector<const void*> ptrs(n);
      for (int i = 0; i < n; i++) ptrs[i] = j->data[i].data();
      uint64_t maxlen = 0;
      for (uint64_t L : j->lens) maxlen = std::max(maxlen, L);
      std::vector<unsigned char> cen(maxlen * elem_size(j->dtype));
      uint64_t clen = 0;
      int med = -1, rounds = 0;
      const int rc = odba_perform_dba(c, j->dtype, ptrs.data(), j->lens.data(), n, j->os, j->oe, j->norm,
                                      j->max_rounds, ODBA_PATH_FULL, 0, cen.data(), &clen, &med, nullptr, &rounds,
                                      nullptr);
      {
        std::lock_guard<std::mutex> lk(mu);
        in_use -= j->footprint;
        running--;
        j->rc = rc;
        if (rc) j->err = odba_last_error(c);
        j->centroid.assign(cen.begin(), cen.begin() + clen * elem_size(j->dtype));
        j->clen = clen;
        j->medoid = med;
        j->rounds = rounds;
        j->data.clear();
        j->data.shrink_to_fit();
        j->t_done = now();
        j->state = rc ? ODBA_JOB_FAILED : ODBA_JOB_DONE;
      }
      cv_done.notify_all();
      cv_admit.notify_all();
    }
  }
};

extern "C" {

int odba_server_create(int device, uint64_t hbm_budget, int slots, odba_server** out) {
  if (!out || slots < 1 || slots > 32) return ODBA_E_INVALID;
  auto* s = new odba_server();
  s->device = device;
  s->slots = slots;
  for (int k = 0; k < slots; k++) {
    odba_ctx* c = nullptr;
    const int rc = odba_ctx_create(device, &c);
    if (rc) {
      for (odba_ctx* x : s->ctxs) odba_ctx_destroy(x);
      delete s;
      return rc;
    }
    s->ctxs.push_back(c);
  }
  int cus = 0, clk = 0;
  uint64_t mem = 0;
  char name[64];
  odba_device_info(s->ctxs[0], &cus, &clk, &mem, name, sizeof(name));
  s->cus = cus > 0 ? cus : 256;
  s->budget = hbm_budget ? std::min<uint64_t>(hbm_budget, mem) : mem / 10 * 8;
  for (odba_ctx* c : s->ctxs) {
    odba_ctx_set_mem_cap(c, s->budget / (uint64_t)slots);
    odba_ctx_set_graphs(c, 0);  // graph replays serialise the slots' streams (measured)
  }
  for (int k = 0; k < slots; k++) s->workers.emplace_back([s, k] { s->worker(k); });
  *out = s;
  return ODBA_OK;
}

int odba_server_destroy(odba_server* s) {
  if (!s) return ODBA_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    for (Job* j : s->queue) {
      j->state = ODBA_JOB_FAILED;
      j->rc = ODBA_E_BUSY;
    }
    s->queue.clear();
  }
  s->cv_admit.notify_all();
  for (auto& t : s->workers) t.join();
  for (odba_ctx* c : s->ctxs) odba_ctx_destroy(c);
  delete s;
  return ODBA_OK;
}

int odba_server_submit(odba_server* s, int dtype, const void* const* seqs, const uint64_t* lens, int n,
                       int use_open_start, int use_open_end, int norm, int max_rounds, uint64_t* job_id) {
  if (!s || !seqs || !lens || n < 2 || !job_id || max_rounds < 1) return ODBA_E_INVALID;
  const size_t es = elem_size(dtype);
  if (!es) return ODBA_E_INVALID;
  auto j = std::make_unique<Job>();
  j->dtype = dtype;
  j->lens.assign(lens, lens + n);
  j->data.resize(n);
  for (int i = 0; i < n; i++) {
    if (!seqs[i] || lens[i] == 0) return ODBA_E_INVALID;
    const unsigned char* p = (const unsigned char*)seqs[i];
    j->data[i].assign(p, p + lens[i] * es);
  }
  j->os = use_open_start;
  j->oe = use_open_end;
  j->norm = norm;
  j->max_rounds = max_rounds;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return ODBA_E_BUSY;
    j->id = s->next_id++;
    j->footprint = s->footprint(*j);
    j->t_submit = s->now();
    *job_id = j->id;
    s->queue.push_back(j.get());
    s->jobs[j->id] = std::move(j);
  }
  s->cv_admit.notify_all();
  return ODBA_OK;
}

int odba_server_wait(odba_server* s, uint64_t job_id, double timeout_s, int* state) {
  if (!s) return ODBA_E_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return ODBA_E_INVALID;
  Job* j = it->second.get();
  auto done = [&] { return j->state == ODBA_JOB_DONE || j->state == ODBA_JOB_FAILED; };
  if (timeout_s < 0) s->cv_done.wait(lk, done);
  else s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
  if (state) *state = j->state;
  return ODBA_OK;
}

int odba_server_result(odba_server* s, uint64_t job_id, void* centroid_out, uint64_t capacity,
                       uint64_t* centroid_len, int* medoid, int* rounds, double* times, int* job_rc) {
  if (!s) return ODBA_E_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return ODBA_E_INVALID;
  Job* j = it->second.get();
  if (j->state != ODBA_JOB_DONE && j->state != ODBA_JOB_FAILED) return ODBA_E_BUSY;
  if (centroid_out) {
    if (capacity < j->clen) return ODBA_E_INVALID;
    std::memcpy(centroid_out, j->centroid.data(), j->centroid.size());
  }
  if (centroid_len) *centroid_len = j->clen;
  if (medoid) *medoid = j->medoid;
  if (rounds) *rounds = j->rounds;
  if (times) {
    times[0] = j->t_submit;
    times[1] = j->t_admit;
    times[2] = j->t_done;
  }
  if (job_rc) *job_rc = j->rc;
  return ODBA_OK;
}

int odba_server_release(odba_server* s, uint64_t job_id) {
  if (!s) return ODBA_E_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return ODBA_E_INVALID;
  if (it->second->state != ODBA_JOB_DONE && it->second->state != ODBA_JOB_FAILED) return ODBA_E_BUSY;
  s->jobs.erase(it);
  return ODBA_OK;
}

int odba_server_stats(odba_server* s, uint64_t* bytes_in_use, uint64_t* budget, int* running, int* queued) {
  if (!s) return ODBA_E_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (bytes_in_use) *bytes_in_use = s->in_use;
  if (budget) *budget = s->budget;
  if (running) *running = s->running;
  if (queued) *queued = (int)s->queue.size();
  return ODBA_OK;
}

}  // extern "C"
