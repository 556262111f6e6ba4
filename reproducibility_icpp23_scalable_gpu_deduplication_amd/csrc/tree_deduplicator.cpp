// tree_deduplicator.cpp -- the C++ TreeDeduplicator facade over the C ABI.
#include "../../include/tree_deduplicator.hpp"

#include <string.h>

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
  const uint32_t id = current_id();
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
