// greedysatostop's on-disk next-state image (FileImage, src/greedysatostop.rs:478-594): the
// packed (n+1)-bit image of all 2^n states in one temporary file, written chunk by chunk in
// state order and then read and rewritten a partition (2^n / partitions states) at a time.
// The bytes are the reference's: little-endian u64 words, entry i at bit i*(n+1).
#pragma once
#include <cstdint>
#include <string>

namespace ns {

class FileImage {
 public:
  // Creates `<prefix>greedy_temp_<nanoseconds>` (the reference's name); partitions is a power
  // of two with partitions * 64 <= 2^state_len; state_len < 64. The file is removed when the
  // object is destroyed.
  FileImage(uint32_t state_len, uint64_t partitions, const std::string& prefix);
  ~FileImage();
  FileImage(const FileImage&) = delete;
  FileImage& operator=(const FileImage&) = delete;

  uint64_t partition_words() const { return partition_bytes_ / 8; }
  uint64_t partition_states() const { return (1ull << state_len_) / partitions_; }
  uint64_t count() const { return count_; }
  const std::string& path() const { return path_; }

  // Appends packed words for the next states (whole entries: words*64 divisible by n+1).
  void save_chunk(const uint64_t* words, uint64_t nwords);
  // Partition part (every state written first): its partition_words() words.
  void load_partition(uint64_t part, uint64_t* words);
  void save_partition(uint64_t part, const uint64_t* words);

 private:
  void check_part(uint64_t part) const;
  uint32_t state_len_;
  uint64_t partitions_;
  int fd_ = -1;
  std::string path_;
  uint64_t partition_bytes_;
  uint64_t count_ = 0;
};

}  // namespace ns
