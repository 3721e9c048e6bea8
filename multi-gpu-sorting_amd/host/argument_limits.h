// Command-line surface of sort_benchmark: the reference's flags, defaults and
// validation (src/argument_limits.cuh:15-311, src/sort_benchmark.cu:16-36),
// with its defects fixed:
//   D5  --num_buffers is validated against {2, 3} (the reference validates it
//       as a merge group size, src/argument_limits.cuh:124-132);
//   D6  GPU counts must be powers of two (the reference's check
//       `size & (size - 1) == 0` parses as `size & ((size - 1) == 0)`, :139,
//       :200);
//   D7  `std` sorts with std::sort (the reference accepts it but has no
//       branch, src/sort_benchmark.cu:73-109).
// Additions: algorithm `hip` (in-HBM sort: inputs copied before the timed
// phase), data types u32/u64, the added distributions of the data generator,
// --values (key-value pairs), --partitions (virtual Tanasic partitions on one
// device), --merge_backend (HET final merge on the CPU or the GPU).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace mgsort {

class ArgumentLimits {
 public:
  ArgumentLimits() = delete;

  static bool ParseNumElements(const std::string& s, size_t* v);
  static bool ParseAlgorithm(const std::string& s, std::string* v);
  static bool ParseGpus(const std::string& s, std::vector<int>* v);
  static bool ParseDataType(const std::string& s, std::string* v);
  static bool ParseDistributionType(const std::string& s, std::string* v);
  static bool ParseDistributionSeed(const std::string& s, uint32_t* v);
  static bool ParseNumThreads(const std::string& s, int* v);
  static bool ParseChunkSize(const std::string& s, size_t* v);
  static bool ParseMergeGroupSize(const std::string& s, size_t* v);
  static bool ParseNumBuffers(const std::string& s, size_t* v);
  static bool ParseValues(const std::string& s, std::string* v);
  static bool ParsePartitions(const std::string& s, size_t* v);
  static bool ParseMergeBackend(const std::string& s, std::string* v);

  static size_t GetDefaultNumElements() { return 0; }
  static std::string GetDefaultAlgorithm() { return ValidAlgorithms().front(); }
  static std::vector<int> GetDefaultGpus();
  static std::string GetDefaultDataType() { return ValidDataTypes().front(); }
  static std::string GetDefaultDistributionType() { return "uniform"; }
  static uint32_t GetDefaultDistributionSeed() { return 2147483647u; }
  static int GetDefaultNumThreads();
  static size_t GetDefaultChunkSize() { return 0; }
  static size_t GetDefaultMergeGroupSize() { return 1; }
  static size_t GetDefaultNumBuffers() { return 2; }
  static std::string GetDefaultValues() { return "none"; }
  static size_t GetDefaultPartitions() { return 0; }  // 0: one per GPU
  static std::string GetDefaultMergeBackend() { return "cpu"; }

  static const std::vector<std::string>& ValidAlgorithms();
  static const std::vector<std::string>& ValidDataTypes();
  static std::vector<std::string> ValidDistributionTypes();
  static const std::vector<int>& ValidGpus();

  static std::string Usage();
};

}  // namespace mgsort
