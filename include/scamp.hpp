// scamp.hpp -- SCAMP's C++ API on the MI355X engine.
//
// The same names, fields and semantics as the reference's public interface,
// so that callers of SCAMP::do_SCAMP compile unchanged against this header:
//   SCAMPArgs                      src/common.h:164-188
//   Profile / ProfileData          src/common.h:98-160
//   mp_entry                       src/common.h:70-74
//   SCAMPProfileType               src/common.h:34-44
//   SCAMPPrecisionType             src/common.h:52-58
//   SCAMPError_t                   src/common.h:359-367
//   SCAMPException                 src/scamp_exception.h:7-22
//   do_SCAMP (two overloads)       src/SCAMP.h:15-18
//   InitProfileMemory, ConvertToEuclidean, CleanupPearson
//                                  src/scamp_utils.h / scamp_utils.cpp:127-141,226-245
// The work runs through the C ABI in scamp_hip.h (one GPU).  ALL_NEIGHBORS
// (approximate KNN) and the multi-device / CPU-worker farm are not part of
// this engine: requesting ALL_NEIGHBORS throws.
#pragma once

#include <cstdint>
#include <exception>
#include <queue>
#include <string>
#include <vector>

class SCAMPException : public std::exception {
 public:
  explicit SCAMPException(std::string message) : _msg(std::move(message)) {}
  const char *what() const noexcept override { return _msg.c_str(); }

 protected:
  std::string _msg;
};

namespace SCAMP {

enum SCAMPProfileType {
  PROFILE_TYPE_INVALID = 0,
  PROFILE_TYPE_1NN_INDEX = 1,
  PROFILE_TYPE_SUM_THRESH = 2,
  PROFILE_TYPE_FREQUENCY_THRESH = 3,
  PROFILE_TYPE_KNN = 4,
  PROFILE_TYPE_1NN_MULTIDIM = 5,
  PROFILE_TYPE_1NN = 6,
  PROFILE_TYPE_APPROX_ALL_NEIGHBORS = 7,
  PROFILE_TYPE_MATRIX_SUMMARY = 8,
};

enum SCAMPPrecisionType {
  PRECISION_INVALID = 0,
  PRECISION_SINGLE = 1,
  PRECISION_MIXED = 2,
  PRECISION_DOUBLE = 3,
  PRECISION_ULTRA = 4,
};

enum SCAMPError_t {
  SCAMP_NO_ERROR,
  SCAMP_FUNCTIONALITY_UNIMPLEMENTED,
  SCAMP_TILE_ILLEGAL_TYPE,
  SCAMP_CUDA_ERROR,
  SCAMP_CUFFT_ERROR,
  SCAMP_CUFFT_EXEC_ERROR,
  SCAMP_DIM_INCOMPATIBLE
};

// Packed 1NN_INDEX entry: floats[0] = FP32 Pearson correlation (low word),
// ints[1] = index (high word).
typedef union {
  float floats[2];
  unsigned int ints[2];
  uint64_t ulong;
} mp_entry;

// One ALL_NEIGHBORS match (src/common.h:76-96): `row` is the matching
// subsequence of subsequence `col`.  Ordered by column, then best first.
struct SCAMPmatch {
  SCAMPmatch() : corr(-2), row(0), col(0) {}
  SCAMPmatch(float d, uint32_t r, uint32_t c) : corr(d), row(r), col(c) {}
  bool operator<(const SCAMPmatch &other) const {
    if (col == other.col) return corr > other.corr;
    return col < other.col;
  }
  float corr;
  uint32_t row;
  uint32_t col;
};

// Heap order of a column's matches: the top is the worst kept match.
class compareMatch {
 public:
  bool operator()(const SCAMPmatch &x1, const SCAMPmatch &x2) {
    return x1.corr > x2.corr;
  }
};

struct ProfileData {
  std::vector<uint32_t> uint32_value;
  std::vector<uint64_t> uint64_value;
  std::vector<float> float_value;
  std::vector<double> double_value;
  // ALL_NEIGHBORS: per subsequence, its kept matches (at most
  // max_matches_per_column)
  std::vector<std::priority_queue<SCAMPmatch, std::vector<SCAMPmatch>, compareMatch>>
      match_value;
};

class Profile {
 public:
  Profile() : type(PROFILE_TYPE_INVALID) {}
  Profile(SCAMPProfileType t, size_t size, float thresh_init = 0,
          int64_t mwidth = -1, int64_t mheight = -1)
      : type(t) {
    Alloc(size, mheight, mwidth, thresh_init);
  }
  std::vector<ProfileData> data;
  std::vector<float> thresholds;
  SCAMPProfileType type;
  // Initial values as SCAMP allocates them (src/common.cpp:283-319).
  void Alloc(size_t size, int64_t matrix_height, int64_t matrix_width,
             float default_thresh);
};

struct SCAMPArgs {
  void validate();
  void print();

  std::vector<double> timeseries_a;
  std::vector<double> timeseries_b;
  Profile profile_a;
  Profile profile_b;
  bool has_b = false;
  uint64_t window = 0;
  uint64_t max_tile_size = 1 << 17;
  int64_t distributed_start_row = -1;
  int64_t distributed_start_col = -1;
  double distance_threshold = 0;
  SCAMPPrecisionType precision_type = PRECISION_DOUBLE;
  SCAMPProfileType profile_type = PROFILE_TYPE_1NN_INDEX;
  bool computing_rows = true;
  bool computing_columns = true;
  bool keep_rows_separate = false;
  bool is_aligned = false;
  bool silent_mode = true;
  int64_t max_matches_per_column = 5;
  int64_t matrix_height = 50;
  int64_t matrix_width = 50;
  // Mixed precision only: rows between FP64 re-anchors (0 = default).  Not
  // a field of the reference's SCAMPArgs; defaulted so that existing callers
  // are unaffected.
  int64_t anchor_interval = 0;
};

// Runs the join on the first visible HIP device (throws SCAMPException when
// there is none).
void do_SCAMP(SCAMPArgs *args);
// `devices` = HIP device ids; the engine runs on devices[0] (a single GPU);
// `num_threads` (the reference's CPU workers) is accepted and unused.
void do_SCAMP(SCAMPArgs *args, const std::vector<int> &devices,
              int num_threads);

std::string getSCAMPErrorString(SCAMPError_t err);

}  // namespace SCAMP

// src/scamp_utils.cpp:127-141, 226-245
bool InitProfileMemory(SCAMP::SCAMPArgs *args);
double ConvertToEuclidean(double val, int window);
double CleanupPearson(double val);

// The file formats of SCAMP's command-line tool (src/scamp_utils.cpp:1-224).
// readFile: one value per whitespace-separated token (std::stod, so "NaN"
// and "inf" parse; blank lines are skipped); a value is kept only when some
// character follows it, so a final value without a line end is dropped, as
// in the reference; an unparsable token prints a FATAL ERROR naming its
// position and exits the process with status 1, as the reference does.
// Returns false (after a message) when the file cannot be opened.
bool readFile(const std::string &filename, std::vector<double> &v);
// Comma-separated integers ("0,1,3"); empty -> empty list.
std::vector<int> ParseIntList(const std::string &s);
SCAMP::SCAMPPrecisionType GetPrecisionType(bool ultrap, bool doublep,
                                           bool mixedp, bool singlep);
SCAMP::SCAMPProfileType ParseProfileType(const std::string &s);
// Writes profile p as text: 1NN_INDEX -> distances (or correlations with
// output_pearson) to `mp` and indices (-1 = no match) to `mpi`, one per line;
// 1NN -> `mp`; SUM_THRESH -> sums; MATRIX_SUMMARY -> matrix_height lines of
// matrix_width space-separated values.  Values with 10 significant digits.
bool WriteProfileToFile(const std::string &mp, const std::string &mpi,
                        SCAMP::Profile &p, bool output_pearson, int window,
                        int matrix_width, int matrix_height);
