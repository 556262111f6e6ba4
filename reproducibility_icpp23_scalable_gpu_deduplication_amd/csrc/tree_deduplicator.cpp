// tree_deduplicator.cpp -- the C++ TreeDeduplicator facade over the C ABI.
#include "../../include/tree_deduplicator.hpp"

#include <string.h>

#include <hip/hip_runtime.h>

#include <fstream>
#include <sstream>
#include <stdexcept>

#include "dd_device.hpp"

namespace dd {

namespace {

void check(dd_status s, const char* what) {
  if (s != DD_OK) throw std::runtime_error(std::string(what) + ": " + dd_strerror(s));
}

HostBuf take(dd_buf& b) {
  HostBuf v(b.data, b.data + b.len);
  dd_buf_free(&b);
  return v;
}

std::pair<uint64_t, uint64_t> sizes_of(int method, const header_t& hd, uint64_t total) {
  if (method == DD_METHOD_LIST) {
    const uint64_t data = (uint64_t)hd.chunk_size * hd.num_first_ocur;
    return {data, total - data};
  }
  const uint64_t meta = 4ull * hd.num_first_ocur + 8ull * hd.num_prior_chkpts + 8ull * hd.num_shift_dupl;
  return {total - meta, meta};
}

}  // namespace

TreeDeduplicator::TreeDeduplicator(uint32_t chunk_size, HashKind kind, void* hip_stream)
    : TreeDeduplicator(chunk_size, kind, hip_stream, DD_METHOD_TREE, "Tree", ".hashtree.incr_chkpt") {}

TreeDeduplicator::TreeDeduplicator(uint32_t chunk_size, HashKind kind, void* hip_stream, int method,
                                   const char* approach, const char* suffix)
    : chunk_size_(chunk_size), method_(method), approach_(approach), suffix_(suffix) {
  check(dd_create_method(chunk_size, static_cast<int>(kind), method, hip_stream, &h_), "dd_create_method");
}

TreeDeduplicator::~TreeDeduplicator() {
  if (h_) dd_destroy(h_);
}

uint32_t TreeDeduplicator::current_id() const {
  uint32_t c = 0;
  dd_get_ids(h_, &c, nullptr);
  return c;
}

uint32_t TreeDeduplicator::baseline_id() const {
  uint32_t b = 0;
  dd_get_ids(h_, nullptr, &b);
  return b;
}

void TreeDeduplicator::set_current_id(uint32_t id) { dd_set_current_id(h_, id); }

void TreeDeduplicator::checkpoint(header_t& header, uint8_t* data_ptr, size_t data_len, HostBuf& diff_h,
                                  bool make_baseline) {
  dd_buf b{nullptr, 0};
  check(dd_checkpoint_header(h_, &header, data_ptr, data_len, make_baseline ? 1 : 0, &b), "checkpoint");
  diff_h = take(b);
  data_len_ = data_len;
  datasizes_ = sizes_of(method_, header, diff_h.size());
}

void TreeDeduplicator::checkpoint(uint8_t* data_ptr, size_t len, std::string& filename, std::string& logname,
                                  bool make_baseline) {
  HostBuf diff_h;
  header_t header;
  checkpoint(header, data_ptr, len, diff_h, make_baseline);
  write_chkpt_log(header, diff_h, logname);
  std::ofstream file;
  file.exceptions(std::ofstream::failbit | std::ofstream::badbit);
  file.open(filename, std::ofstream::out | std::ofstream::binary);
  file.write(reinterpret_cast<const char*>(diff_h.data()), (std::streamsize)diff_h.size());
  file.close();
  set_current_id(current_id() + 1);
}

void TreeDeduplicator::checkpoint(uint8_t* data_ptr, size_t len, HostBuf& diff_h, bool make_baseline) {
  header_t header;
  checkpoint(header, data_ptr, len, diff_h, make_baseline);
  set_current_id(current_id() + 1);
}

void TreeDeduplicator::checkpoint(uint8_t* data_ptr, size_t len, HostBuf& diff_h, std::string& logname,
                                  bool make_baseline) {
  header_t header;
  checkpoint(header, data_ptr, len, diff_h, make_baseline);
  write_chkpt_log(header, diff_h, logname);
  set_current_id(current_id() + 1);
}

void TreeDeduplicator::restart(uint8_t* data_ptr, size_t len, std::vector<HostBuf>& chkpts, std::string& logname,
                               uint32_t chkpt_id) {
  std::vector<dd_buf> chain(chkpts.size());
  for (size_t i = 0; i < chkpts.size(); i++) chain[i] = dd_buf{chkpts[i].data(), chkpts[i].size()};
  check(dd_restart(h_, chain.data(), (uint32_t)chain.size(), chkpt_id, data_ptr, len), "restart");
  std::string name = logname + ".chunk_size." + std::to_string(chunk_size_) + ".restart_timing.csv";
  write_restart_log(chkpt_id, name);
}

void TreeDeduplicator::restart(uint8_t* data_ptr, size_t len, std::vector<std::string>& chkpt_filenames,
                               std::string& logname, uint32_t chkpt_id) {
  std::vector<HostBuf> bufs(chkpt_filenames.size());
  for (size_t i = 0; i < chkpt_filenames.size(); i++) {
    std::ifstream f;
    f.exceptions(std::ifstream::failbit | std::ifstream::badbit);
    f.open(chkpt_filenames[i] + suffix_, std::ifstream::in | std::ifstream::binary);
    f.seekg(0, std::ios::end);
    const std::streamoff n = f.tellg();
    f.seekg(0, std::ios::beg);
    bufs[i].resize((size_t)n);
    f.read(reinterpret_cast<char*>(bufs[i].data()), n);
  }
  std::vector<dd_buf> chain(bufs.size());
  for (size_t i = 0; i < bufs.size(); i++) chain[i] = dd_buf{bufs[i].data(), bufs[i].size()};
  check(dd_restart(h_, chain.data(), (uint32_t)chain.size(), chkpt_id, data_ptr, len), "restart");
  write_restart_log(chkpt_id, logname);
}

void TreeDeduplicator::checkpoint_batch(const std::vector<TreeDeduplicator*>& dedups,
                                        const std::vector<uint8_t*>& data, const std::vector<size_t>& lens,
                                        std::vector<std::string>& filenames, std::vector<std::string>& lognames,
                                        bool make_baseline) {
  const size_t n = dedups.size();
  if (n == 0) return;
  if (data.size() != n || lens.size() != n || filenames.size() != n || lognames.size() != n)
    throw std::runtime_error("checkpoint_batch: argument sizes differ");
  std::vector<dd_dedup*> hs(n);
  for (size_t i = 0; i < n; i++) hs[i] = dedups[i]->h_;
  dd_batch* b = nullptr;
  check(dd_batch_create(hs.data(), (uint32_t)n, &b), "dd_batch_create");
  // Buffer shapes change from call to call here: no graph capture.
  dd_batch_use_graph(b, 0);
  std::vector<void*> outs(n, nullptr);
  std::vector<uint64_t> caps(n), out_lens(n), ulens(n);
  std::vector<int> base(n, make_baseline ? 1 : 0);
  std::vector<const void*> ins(n);
  dd_status st = DD_OK;
  for (size_t i = 0; i < n && st == DD_OK; i++) {
    caps[i] = dd_max_diff_size(lens[i], dedups[i]->chunk_size_, dedups[i]->current_id());
    ulens[i] = lens[i];
    ins[i] = data[i];
    if (hipMalloc(&outs[i], caps[i]) != hipSuccess) st = DD_ENOMEM;
  }
  if (st == DD_OK)
    st = dd_batch_checkpoint_device(b, ins.data(), ulens.data(), base.data(), outs.data(), caps.data(), out_lens.data());
  std::vector<HostBuf> diffs(n);
  for (size_t i = 0; i < n && st == DD_OK; i++) {
    diffs[i].resize(out_lens[i]);
    if (hipMemcpy(diffs[i].data(), outs[i], out_lens[i], hipMemcpyDeviceToHost) != hipSuccess) st = DD_EHIP;
  }
  for (void* p : outs)
    if (p) (void)hipFree(p);
  dd_batch_destroy(b);
  check(st, "checkpoint_batch");
  for (size_t i = 0; i < n; i++) {
    TreeDeduplicator& d = *dedups[i];
    header_t header;
    memcpy(&header, diffs[i].data(), sizeof header);
    d.data_len_ = lens[i];
    d.datasizes_ = sizes_of(d.method_, header, diffs[i].size());
    d.write_chkpt_log(header, diffs[i], lognames[i]);
    std::ofstream file;
    file.exceptions(std::ofstream::failbit | std::ofstream::badbit);
    file.open(filenames[i], std::ofstream::out | std::ofstream::binary);
    file.write(reinterpret_cast<const char*>(diffs[i].data()), (std::streamsize)diffs[i].size());
    file.close();
  }
}

size_t TreeDeduplicator::num_first_ocur() {
  uint64_t f = 0;
  dd_counts(h_, &f, nullptr);
  return (size_t)f;
}

size_t TreeDeduplicator::num_shift_dupl() {
  uint64_t s = 0;
  dd_counts(h_, nullptr, &s);
  return (size_t)s;
}

std::string metadata_breakdown(const header_t& header, const HostBuf& diff, uint32_t num_chkpts) {
  std::ostringstream fs;
  fs << 32 << "," << 4ull * header.num_first_ocur << ",";
  const uint8_t* table = diff.data() + 32 + 4ull * header.num_first_ocur;
  if (header.ref_id != header.chkpt_id) {
    fs << 8ull * header.num_prior_chkpts;
    for (uint32_t i = 0; i < num_chkpts; i++) {
      uint64_t bytes = 0;
      if (i < header.num_prior_chkpts && diff.size() >= 32 + 4ull * header.num_first_ocur + 8ull * (i + 1)) {
        uint32_t num;
        memcpy(&num, table + 8ull * i + 4, 4);
        bytes = 8ull * num;
      }
      fs << "," << bytes;
    }
  } else {
    fs << 0 << "," << 8ull * header.num_shift_dupl;
    for (uint32_t i = 1; i < num_chkpts; i++) fs << "," << 0;
  }
  fs << "\n";
  return fs.str();
}

void TreeDeduplicator::write_chkpt_log(header_t& header, HostBuf& diff_h, std::string& logname) {
  double t[4], rt[2];
  dd_get_timers(h_, t, rt);
  const std::string base = logname + ".chunk_size." + std::to_string(chunk_size_);
  std::fstream result(base + ".csv", std::fstream::out | std::fstream::app);
  std::fstream size_file(base + ".size.csv", std::fstream::out | std::fstream::app);
  std::fstream timing(base + ".timing.csv", std::fstream::out | std::fstream::app);
  result.seekp(0, std::ios::end);
  if (result.tellp() == 0)
    result << "Approach,Chkpt ID,Chunk Size,Uncompressed Size,Compressed Size,Data Size,Metadata Size,"
              "Setup Time,Comparison Time,Gather Time,Write Time\n";
  const uint32_t id = header.chkpt_id;  // the id of the checkpoint being logged
  result << approach_ << "," << id << "," << chunk_size_ << "," << data_len_ << "," << datasizes_.first + datasizes_.second
         << "," << datasizes_.first << "," << datasizes_.second << "," << t[0] << "," << t[1] << "," << t[2] << ","
         << t[3] << "\n";
  timing << approach_ << "," << id << "," << chunk_size_ << "," << t[1] << "," << t[2] << "," << t[3] << "\n";
  size_file << approach_ << "," << id << "," << chunk_size_ << "," << datasizes_.first << "," << datasizes_.second << ","
            << metadata_breakdown(header, diff_h, id);
}

void TreeDeduplicator::write_restart_log(uint32_t select_chkpt, std::string& logname) {
  double t[4], rt[2];
  dd_get_timers(h_, t, rt);
  std::fstream f(logname, std::fstream::out | std::fstream::app);
  f << approach_ << "," << select_chkpt << "," << chunk_size_ << "," << rt[0] << "," << rt[1] << "\n";
}

}  // namespace dd
