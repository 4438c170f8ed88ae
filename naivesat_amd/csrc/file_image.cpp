#include "file_image.h"

#include <fcntl.h>
#include <unistd.h>

#include <cerrno>
#include <chrono>
#include <cstring>

#include "runtime.h"

namespace ns {

namespace {

void io_all(int fd, void* buf, uint64_t bytes, uint64_t off, bool write, const std::string& path) {
  char* p = static_cast<char*>(buf);
  while (bytes) {
    const size_t step = bytes > (1ull << 30) ? (1ull << 30) : (size_t)bytes;
    const ssize_t k = write ? ::pwrite(fd, p, step, (off_t)off) : ::pread(fd, p, step, (off_t)off);
    if (k < 0 && errno == EINTR) continue;
    if (k <= 0)
      throw Error(ST_INVALID, std::string(write ? "write to " : "read from ") + path + ": " +
                                  (k < 0 ? std::strerror(errno) : "unexpected end of file"));
    p += k;
    off += (uint64_t)k;
    bytes -= (uint64_t)k;
  }
}

}  // namespace

FileImage::FileImage(uint32_t state_len, uint64_t partitions, const std::string& prefix)
    : state_len_(state_len), partitions_(partitions) {
  if (state_len >= 64) throw Error(ST_INVALID, "file images take fewer than 64 state bits");
  if (partitions == 0 || (partitions & (partitions - 1)))
    throw Error(ST_INVALID, "the number of partitions must be a power of two");
  if (partitions * 64 > (1ull << state_len)) throw Error(ST_INVALID, "at most 2^n / 64 partitions");
  const auto ns = std::chrono::duration_cast<std::chrono::nanoseconds>(
                      std::chrono::system_clock::now().time_since_epoch())
                      .count();
  path_ = prefix + "greedy_temp_" + std::to_string(ns);
  fd_ = ::open(path_.c_str(), O_RDWR | O_CREAT, 0644);
  if (fd_ < 0) throw Error(ST_INVALID, "cannot create " + path_ + ": " + std::strerror(errno));
  partition_bytes_ = (((1ull << state_len) / partitions) * (state_len + 1)) >> 3;
}

FileImage::~FileImage() {
  if (fd_ >= 0) ::close(fd_);
  ::unlink(path_.c_str());
}

void FileImage::save_chunk(const uint64_t* words, uint64_t nwords) {
  const uint64_t bits = nwords * 64;
  if (bits % (state_len_ + 1)) throw Error(ST_INVALID, "a chunk holds whole entries");
  const uint64_t add = bits / (state_len_ + 1);
  if (count_ + add > (1ull << state_len_)) throw Error(ST_INVALID, "more entries than states");
  io_all(fd_, const_cast<uint64_t*>(words), nwords * 8, (count_ * (state_len_ + 1)) >> 3, true, path_);
  count_ += add;
}

void FileImage::check_part(uint64_t part) const {
  if (part >= partitions_) throw Error(ST_INVALID, "partition out of range");
  if (count_ != (1ull << state_len_)) throw Error(ST_INVALID, "the image is not complete");
}

void FileImage::load_partition(uint64_t part, uint64_t* words) {
  check_part(part);
  io_all(fd_, words, partition_bytes_, part * partition_bytes_, false, path_);
}

void FileImage::save_partition(uint64_t part, const uint64_t* words) {
  check_part(part);
  io_all(fd_, const_cast<uint64_t*>(words), partition_bytes_, part * partition_bytes_, true, path_);
}

}  // namespace ns
