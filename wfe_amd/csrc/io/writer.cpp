// Derived in part from Xassemblianist/WFE, Copyright (c) 2026 Omer Kaan (Xassemblianist),
// MIT License; the licence text is in THIRD_PARTY_NOTICES at the repository root.
#include "io/writer.hpp"

#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <stdexcept>

namespace wfe {

Writer::Writer() {
  if (const char* e = std::getenv("WFE_SYNC_OUTPUT")) async_ = std::string(e) != "1";
}

Writer::~Writer() {
  try {
    finish();
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
  }
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  if (thread_.joinable()) thread_.join();
  for (float*& p : slot_mem_)
    if (p) (void)wfe_host_free(p);
}

void Writer::init(const GDims& g, const std::string& out_dir, double dt, bool moisture,
                  bool has_tsk) {
  g_ = g;
  dir_ = out_dir;
  moist_ = moisture;
  sfc_ = has_tsk;
  std::filesystem::create_directories(dir_);
  const std::string meta = dir_ + "/meta.json";
  std::FILE* f = std::fopen(meta.c_str(), "w");
  if (!f) throw std::runtime_error("cannot write " + meta);
  std::fprintf(f,
               "{\n"
               "  \"nx\": %d, \"ny\": %d, \"nz\": %d,\n"
               "  \"dx\": %g, \"dy\": %g, \"dz\": %g,\n"
               "  \"dt\": %g,\n"
               "  \"vars\": {\"u\": %d, \"v\": %d, \"w\": %d, \"thp\": %d, \"pip\": %d",
               g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, dt, g.nz, g.nz, g.nz + 1, g.nz, g.nz);
  if (moist_)
    std::fprintf(f, ", \"qv\": %d, \"qc\": %d, \"qr\": %d, \"rain\": 1", g.nz, g.nz, g.nz);
  if (has_tsk) std::fprintf(f, ", \"tsk\": 1, \"pblh\": 1, \"t2m\": 1, \"u10\": 1");
  std::fprintf(f, "}\n}\n");
  std::fclose(f);

  per_field_ = (size_t)g.nx * g.ny * (g.nz + 1);
  if (async_) {
    for (int s = 0; s < kSlots; ++s) {
      void* p = nullptr;
      if (wfe_host_alloc(per_field_ * kFieldsPerSlot * sizeof(float), &p) != WFE_OK)
        throw std::runtime_error(std::string("pinned output buffer: ") + wfe_last_error());
      slot_mem_[s] = static_cast<float*>(p);
    }
    thread_ = std::thread([this] { worker(); });
  } else {
    full_.resize(g.npts());
    out_.resize(per_field_);
  }
}

std::vector<Writer::Item> Writer::items_for(bool with_2d) const {
  const size_t n2 = (size_t)g_.nx * g_.ny;
  const size_t n3 = n2 * g_.nz;
  std::vector<Item> it = {{"u", WFE_F_U, n3},
                          {"v", WFE_F_V, n3},
                          {"w", WFE_F_W, n3 + n2},
                          {"thp", WFE_F_THP, n3},
                          {"pip", WFE_F_PIP, n3}};
  if (moist_) {
    it.push_back({"qv", WFE_F_QV, n3});
    it.push_back({"qc", WFE_F_QC, n3});
    it.push_back({"qr", WFE_F_QR, n3});
    if (with_2d) it.push_back({"rain", WFE_F_RAIN, n2});
  }
  if (with_2d && sfc_) {  // src/main.cpp:460-465
    it.push_back({"tsk", WFE_F_TSK, n2});
    it.push_back({"pblh", WFE_F_PBLH, n2});
    it.push_back({"t2m", WFE_F_T2M, n2});
    it.push_back({"u10", WFE_F_U10, n2});
  }
  return it;
}

std::string Writer::path_of(const std::string& name, int step) const {
  char path[1024];
  std::snprintf(path, sizeof(path), "%s/%s_%06d.bin", dir_.c_str(), name.c_str(), step);
  return path;
}

void Writer::write_raw(const std::string& path, const float* data, size_t n) {
  std::FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot open output file " + path);
  const size_t wrote = std::fwrite(data, sizeof(float), n, f);
  std::fclose(f);
  if (wrote != n) throw std::runtime_error("short write to " + path);
}

// The reference's path (src/io/writer.cpp:89-107): whole-field download,
// interior packed on the host.
void Writer::write_sync(wfe_ctx* ctx, int step, const std::vector<Item>& items) {
  for (const Item& it : items) {
    const bool two_d = it.field >= WFE_F_RAIN;
    const size_t nfull = two_d ? g_.n2() : g_.npts();
    if (wfe_download(ctx, it.field, full_.data(), nfull) != WFE_OK)
      throw std::runtime_error(std::string("download failed: ") + wfe_last_error());
    const int nlev = two_d ? 1 : (it.field == WFE_F_W ? g_.nz + 1 : g_.nz);
    size_t o = 0;
    for (int k = 0; k < nlev; ++k)
      for (int j = 0; j < g_.ny; ++j)
        for (int i = 0; i < g_.nx; ++i)
          out_[o++] = (float)full_[two_d ? g_.idx2(i, j) : g_.idx(i, j, k)];
    write_raw(path_of(it.name, step), out_.data(), o);
  }
}

void Writer::rethrow_locked() {
  if (err_) {
    std::exception_ptr e = err_;
    err_ = nullptr;
    std::rethrow_exception(e);
  }
}

void Writer::write(wfe_ctx* ctx, int step, double t, bool with_2d) {
  const std::vector<Item> items = items_for(with_2d);
  if (!async_) {
    write_sync(ctx, step, items);
  } else {
    int slot;
    {
      std::unique_lock<std::mutex> lk(mu_);
      rethrow_locked();
      slot = next_slot_;
      cv_.wait(lk, [&] { return !slot_busy_[slot] || err_; });
      rethrow_locked();
      slot_busy_[slot] = true;
      next_slot_ = (slot + 1) % kSlots;
    }
    Job job{ctx, slot, step, 0, items};
    try {
      for (size_t q = 0; q < items.size(); ++q) {
        float* dst = slot_mem_[slot] + q * per_field_;
        if (wfe_pack_field_async(ctx, items[q].field, dst, items[q].n) != WFE_OK)
          throw std::runtime_error(std::string("output pack failed: ") + wfe_last_error());
      }
      if (wfe_output_mark(ctx, &job.ticket) != WFE_OK)
        throw std::runtime_error(std::string("output mark failed: ") + wfe_last_error());
    } catch (...) {
      std::lock_guard<std::mutex> lk(mu_);
      slot_busy_[slot] = false;
      throw;
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      queue_.push_back(std::move(job));
    }
    cv_.notify_all();
  }
  std::printf("  cikti yazildi: step %d (t = %.1f s)\n", step, t);
}

void Writer::worker() {
  for (;;) {
    Job job;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
      if (queue_.empty()) return;
      job = std::move(queue_.front());
      queue_.pop_front();
    }
    try {
      if (wfe_output_wait(job.ctx, job.ticket) != WFE_OK)
        throw std::runtime_error(std::string("output copy failed: ") + wfe_last_error());
      for (size_t q = 0; q < job.items.size(); ++q)
        write_raw(path_of(job.items[q].name, job.step), slot_mem_[job.slot] + q * per_field_,
                  job.items[q].n);
    } catch (...) {
      std::lock_guard<std::mutex> lk(mu_);
      if (!err_) err_ = std::current_exception();
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      slot_busy_[job.slot] = false;
    }
    cv_.notify_all();
  }
}

void Writer::finish() {
  if (!async_) return;
  std::unique_lock<std::mutex> lk(mu_);
  cv_.wait(lk, [&] {
    if (!queue_.empty()) return false;
    for (bool b : slot_busy_)
      if (b) return false;
    return true;
  });
  rethrow_locked();
}

void Writer::write_static(const char* name, const std::vector<float>& data) {
  char path[1024];
  std::snprintf(path, sizeof(path), "%s/%s.bin", dir_.c_str(), name);
  write_raw(path, data.data(), data.size());
}

}  // namespace wfe
