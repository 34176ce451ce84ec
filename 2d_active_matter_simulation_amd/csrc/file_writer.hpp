// Background file writer of the drop-in CLI (cli.cpp): energy lines and
// snapshots written in submission order on one thread while the GPU steps.
// Host-only (no HIP): the CLI and the ThreadSanitizer check of the host code
// include it.
#pragma once
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/actflow.h"

namespace af_cli {

// %g, the default ostream format of the reference's energy lines
inline std::string fmt(double v) {
  char b[64];
  std::snprintf(b, sizeof b, "%g", v);
  return b;
}

// The files of a run are written by one background thread, in the order the
// jobs are submitted, while the GPU steps the next chunk: energy lines
// (appended to energy.txt, .cu:1268-1278) and snapshots (.cu:1245-1266).
// After a job fails no later job is written, so the files on disk are those
// of a run that stopped at that job, as the reference does; the main loop
// sees the failure at its next chunk and exits with the job's message.
class Writer {
 public:
  Writer() : th_([this] { loop(); }) {}
  ~Writer() { finish(); }

  // the energies of steps m .. m + n - 1 (step m labelled m*dt)
  void energy(const std::string& path, long long m, float dt, const double* e, int n) {
    Job j;
    j.snapshot = false;
    j.path = path;
    j.m = m;
    j.dt = dt;
    j.data.assign(e, e + n);
    push(std::move(j));
  }
  // waits while two snapshots are queued (bounds the memory held)
  void snapshot(const std::string& path, int nx, int ny, const std::vector<double>& w, const double M[2],
                int precision) {
    std::unique_lock<std::mutex> lk(mu_);
    space_.wait(lk, [&] { return queued_snaps_ < 2 || failed_; });
    lk.unlock();
    Job j;
    j.snapshot = true;
    j.path = path;
    j.nx = nx;
    j.ny = ny;
    j.precision = precision;
    j.M[0] = M[0];
    j.M[1] = M[1];
    j.data = w;
    push(std::move(j));
  }
  // the message of the first failed job, or "" if none failed so far
  std::string failure() {
    std::lock_guard<std::mutex> lk(mu_);
    return msg_;
  }
  // writes everything queued, stops the thread; returns failure()
  std::string finish() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      done_ = true;
    }
    work_.notify_one();
    if (th_.joinable()) th_.join();
    return failure();
  }

 private:
  struct Job {
    bool snapshot = false;
    std::string path;
    long long m = 0;
    float dt = 0;
    int nx = 0, ny = 0, precision = 6;
    double M[2] = {0, 0};
    std::vector<double> data;
  };

  void push(Job&& j) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (failed_) return;
      if (j.snapshot) ++queued_snaps_;
      q_.push_back(std::move(j));
    }
    work_.notify_one();
  }

  void loop() {
    for (;;) {
      Job j;
      {
        std::unique_lock<std::mutex> lk(mu_);
        work_.wait(lk, [&] { return !q_.empty() || done_; });
        if (q_.empty()) return;
        j = std::move(q_.front());
        q_.pop_front();
      }
      const std::string err = j.snapshot ? write_snapshot(j) : write_energy(j);
      std::lock_guard<std::mutex> lk(mu_);
      if (j.snapshot) --queued_snaps_;
      if (!err.empty() && !failed_) {
        failed_ = true;
        msg_ = err;
        for (auto& x : q_)
          if (x.snapshot) --queued_snaps_;
        q_.clear();
      }
      space_.notify_all();
    }
  }

  static std::string write_energy(const Job& j) {
    FILE* ef = std::fopen(j.path.c_str(), "a");
    if (!ef) return "Error: failed to save the mean energy";
    for (size_t i = 0; i < j.data.size(); ++i)
      std::fprintf(ef, "%s %s\n", fmt((float)(j.m + (long long)i) * j.dt).c_str(), fmt(j.data[i]).c_str());
    return std::fclose(ef) == 0 ? "" : "Error: failed to save the mean energy";
  }

  static std::string write_snapshot(const Job& j) {
    return af_snapshot_write(j.path.c_str(), j.nx, j.ny, j.data.data(), j.M, j.precision) == AF_OK
               ? ""
               : "Error: unable to create snapshot file";
  }

  std::mutex mu_;
  std::condition_variable work_, space_;
  std::deque<Job> q_;
  int queued_snaps_ = 0;
  bool done_ = false, failed_ = false;
  std::string msg_;
  std::thread th_;
};

}  // namespace af_cli
