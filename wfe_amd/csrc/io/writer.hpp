#pragma once
// Derived in part from Xassemblianist/WFE, Copyright (c) 2026 Omer Kaan (Xassemblianist),
// MIT License; the licence text is in THIRD_PARTY_NOTICES at the repository root.

#include <condition_variable>
#include <deque>
#include <exception>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "core/grid.hpp"
#include "wfe_hip.h"

namespace wfe {

// Snapshot writer with the reference's on-disk format (src/io/writer.hpp:14-33,
// src/io/writer.cpp:12-107): out_dir/meta.json plus one raw float32 file per
// field and output step, <var>_%06d.bin, interior only, i fastest, k slowest,
// nz levels (nz+1 for w); 2-D fields (rain) likewise; static fields as
// <name>.bin.
//
// Output is asynchronous: write() packs each field's interior as float32 on
// the device (wfe_pack_field_async), the copies run on the context's copy
// stream into one of two pinned host slots, and a writer thread waits for
// them and writes the files while the caller keeps stepping. A third output
// waits until the older of the two slots has been written. Errors of the
// writer thread are rethrown by the next write() or by finish(). With
// synchronous output (the constructor flag, or WFE_SYNC_OUTPUT=1 in the
// environment) every field is downloaded and written before write() returns,
// as the reference does.
class Writer {
 public:
  Writer();
  ~Writer();
  Writer(const Writer&) = delete;
  Writer& operator=(const Writer&) = delete;

  // Throws std::runtime_error when meta.json cannot be written.
  void init(const GDims& g, const std::string& out_dir, double dt, bool moisture, bool has_tsk);
  // The state snapshot of a step; with_2d adds the 2-D fields of an output
  // step: rain (moist runs) and tsk, pblh, t2m, u10 (surface physics).
  void write(wfe_ctx* ctx, int step, double t, bool with_2d = false);
  void write_static(const char* name, const std::vector<float>& data);
  // Waits until every queued snapshot is on disk; rethrows a writer error.
  void finish();
  bool asynchronous() const { return async_; }
  const std::string& dir() const { return dir_; }

 private:
  struct Item {
    std::string name;
    int field;
    size_t n;
  };
  struct Job {
    wfe_ctx* ctx;
    int slot;
    int step;
    long long ticket;
    std::vector<Item> items;
  };
  std::vector<Item> items_for(bool with_2d) const;
  void write_sync(wfe_ctx* ctx, int step, const std::vector<Item>& items);
  void write_raw(const std::string& path, const float* data, size_t n);
  std::string path_of(const std::string& name, int step) const;
  void worker();
  void rethrow_locked();

  GDims g_{};
  std::string dir_;
  bool moist_ = false;
  bool sfc_ = false;
  bool async_ = true;
  std::vector<double> full_;
  std::vector<float> out_;

  // Two pinned slots of nfield_slots_ arrays of nx*ny*(nz+1) floats each.
  static constexpr int kSlots = 2;
  static constexpr int kFieldsPerSlot = WFE_F_COUNT;
  float* slot_mem_[kSlots]{};
  bool slot_busy_[kSlots]{};
  int next_slot_ = 0;
  size_t per_field_ = 0;

  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<Job> queue_;
  bool stop_ = false;
  std::exception_ptr err_;
  std::thread thread_;
};

}  // namespace wfe
