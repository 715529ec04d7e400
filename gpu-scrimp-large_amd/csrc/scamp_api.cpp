// SCAMP's C++ API (include/scamp.hpp) over the engine's C ABI.
//
// do_SCAMP keeps the reference's contract (src/SCAMP.cpp:261-334): validate
// the arguments, run the join, leave the profile(s) in args->profile_a / _b
// (which the caller pre-allocated, e.g. with InitProfileMemory, and which are
// read as the prior best-so-far), throw SCAMPException on failure.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "scamp.hpp"
#include "scamp_hip.h"

namespace SCAMP {

void Profile::Alloc(size_t size, int64_t matrix_height, int64_t matrix_width,
                    float default_thresh) {
  switch (type) {
    case PROFILE_TYPE_SUM_THRESH:
      data.emplace_back();
      data[0].double_value.resize(size, 0);
      break;
    case PROFILE_TYPE_1NN:
      data.emplace_back();
      data[0].float_value.resize(size, -2.0);
      break;
    case PROFILE_TYPE_1NN_INDEX: {
      mp_entry e;
      e.ints[1] = -1u;
      e.floats[0] = -2.0;
      data.emplace_back();
      data[0].uint64_value.resize(size, e.ulong);
      break;
    }
    case PROFILE_TYPE_FREQUENCY_THRESH:
      data.emplace_back();
      data[0].uint64_value.resize(size, 0);
      break;
    case PROFILE_TYPE_APPROX_ALL_NEIGHBORS:
      data.emplace_back();
      data[0].match_value.resize(size);
      thresholds.resize(size, default_thresh);
      break;
    case PROFILE_TYPE_MATRIX_SUMMARY:
      data.emplace_back();
      data[0].float_value.resize(matrix_height * matrix_width, -2.0);
      break;
    default:
      break;
  }
}

// src/common.cpp:417-436 (same checks, same messages)
void SCAMPArgs::validate() {
  if (window < 3) {
    throw SCAMPException("Error: Subsequence length must be at least 3");
  }
  if (max_tile_size < 1024) {
    throw SCAMPException("Error: max tile size must be at least 1024");
  }
  if (max_tile_size / 2 < window) {
    throw SCAMPException(
        "Error: Tile length and width must be at least 2x larger than the "
        "window size");
  }
  if (timeseries_a.size() < window || (has_b && timeseries_b.size() < window)) {
    throw SCAMPException(
        "Error: Input time series must be at least 'subesequence window "
        "size' in length");
  }
}

void SCAMPArgs::print() {
  std::cout << "window: " << window << std::endl;
  std::cout << "max_tile_size: " << max_tile_size << std::endl;
  std::cout << "has_b: " << has_b << std::endl;
  std::cout << "keep_rows_separate: " << keep_rows_separate << std::endl;
  std::cout << "computing_rows: " << computing_rows << std::endl;
  std::cout << "computing_columns: " << computing_columns << std::endl;
  std::cout << "is_aligned: " << is_aligned << std::endl;
  std::cout << "profile_type: " << profile_type << std::endl;
  std::cout << "precision_type: " << precision_type << std::endl;
  std::cout << "distance_threshold: " << distance_threshold << std::endl;
}

std::string getSCAMPErrorString(SCAMPError_t err) {
  switch (err) {
    case SCAMP_NO_ERROR: return "SCAMP_NO_ERROR";
    case SCAMP_FUNCTIONALITY_UNIMPLEMENTED:
      return "SCAMP_FUNCTIONALITY_UNIMPLEMENTED";
    case SCAMP_TILE_ILLEGAL_TYPE: return "SCAMP_TILE_ILLEGAL_TYPE";
    case SCAMP_CUDA_ERROR: return "SCAMP_CUDA_ERROR";
    case SCAMP_CUFFT_ERROR: return "SCAMP_CUFFT_ERROR";
    case SCAMP_CUFFT_EXEC_ERROR: return "SCAMP_CUFFT_EXEC_ERROR";
    case SCAMP_DIM_INCOMPATIBLE: return "SCAMP_DIM_INCOMPATIBLE";
  }
  return "SCAMP_UNKNOWN_ERROR";
}

namespace {

// The device state of one do_SCAMP call: a context (stream + device blocks)
// created on entry and destroyed on every exit path, so the call leaves no
// device memory behind and concurrent calls from different host threads run
// on their own streams, as the reference's per-call workers and tiles do
// (src/SCAMP.cpp:240-256, src/common.cpp:109-124).
class CallContext {
 public:
  explicit CallContext(int device) {
    if (device < 0) throw SCAMPException("Error: invalid device id");
    if (scamp_hip_ctx_create(device, &ctx_) != SCAMP_HIP_OK) {
      char buf[512];
      scamp_hip_last_error(buf, sizeof(buf));
      throw SCAMPException(std::string("Error: no HIP device usable: ") + buf);
    }
  }
  ~CallContext() { scamp_hip_ctx_destroy(ctx_); }
  CallContext(const CallContext &) = delete;
  CallContext &operator=(const CallContext &) = delete;
  scamp_hip_ctx *get() const { return ctx_; }

 private:
  scamp_hip_ctx *ctx_ = nullptr;
};

void *profile_buffer(Profile &p, SCAMPProfileType t) {
  if (p.data.empty()) return nullptr;
  switch (t) {
    case PROFILE_TYPE_1NN_INDEX: return p.data[0].uint64_value.data();
    case PROFILE_TYPE_SUM_THRESH: return p.data[0].double_value.data();
    case PROFILE_TYPE_1NN:
    case PROFILE_TYPE_MATRIX_SUMMARY: return p.data[0].float_value.data();
    default: return nullptr;
  }
}

size_t profile_length(const Profile &p, SCAMPProfileType t) {
  if (p.data.empty()) return 0;
  switch (t) {
    case PROFILE_TYPE_1NN_INDEX: return p.data[0].uint64_value.size();
    case PROFILE_TYPE_SUM_THRESH: return p.data[0].double_value.size();
    case PROFILE_TYPE_APPROX_ALL_NEIGHBORS: return p.data[0].match_value.size();
    default: return p.data[0].float_value.size();
  }
}

// ALL_NEIGHBORS: the per-subsequence queues <-> the C ABI's k-entry arrays.
std::vector<scamp_hip_match> queues_to_matches(Profile &p, size_t n, int64_t k) {
  std::vector<scamp_hip_match> out(n * k);
  for (size_t j = 0; j < n; ++j) {
    auto q = p.data[0].match_value[j];  // copy: the caller's queue stays
    std::vector<SCAMPmatch> v;
    while (!q.empty()) {
      v.push_back(q.top());
      q.pop();
    }
    // best first, at most k of them
    std::vector<scamp_hip_match> row(k, scamp_hip_match{-2.0f, 0xFFFFFFFFu, (uint32_t)j});
    int64_t t = 0;
    for (auto it = v.rbegin(); it != v.rend() && t < k; ++it, ++t)
      row[t] = scamp_hip_match{it->corr, it->row, (uint32_t)j};
    std::copy(row.begin(), row.end(), out.begin() + j * k);
  }
  return out;
}

void matches_to_queues(const std::vector<scamp_hip_match> &m, Profile &p, size_t n,
                       int64_t k) {
  for (size_t j = 0; j < n; ++j) {
    auto &q = p.data[0].match_value[j];
    q = std::priority_queue<SCAMPmatch, std::vector<SCAMPmatch>, compareMatch>();
    for (int64_t t = 0; t < k; ++t) {
      const scamp_hip_match &e = m[j * k + t];
      if (e.corr > -2.0f) q.emplace(e.corr, e.row, (uint32_t)j);
    }
  }
}

}  // namespace

void do_SCAMP(SCAMPArgs *args, const std::vector<int> &devices,
              int num_threads) {
  (void)num_threads;
  if (args == nullptr) {
    throw SCAMPException("Error: Invalid arguments provided to SCAMP");
  }
  args->validate();
  int ndev = 0;
  scamp_hip_device_count(&ndev);
  if (ndev == 0) {
    throw SCAMPException(
        "Error: no compute_resources provided (no HIP device is visible; "
        "this engine has no CPU path)");
  }
  const SCAMPProfileType t = args->profile_type;
  if (t != PROFILE_TYPE_1NN_INDEX && t != PROFILE_TYPE_1NN &&
      t != PROFILE_TYPE_SUM_THRESH && t != PROFILE_TYPE_MATRIX_SUMMARY &&
      t != PROFILE_TYPE_APPROX_ALL_NEIGHBORS) {
    throw SCAMPException("ERROR " +
                         getSCAMPErrorString(SCAMP_FUNCTIONALITY_UNIMPLEMENTED) +
                         " executing tile");
  }
  const int device = devices.empty() ? 0 : devices[0];

  const size_t na = args->timeseries_a.size();
  const size_t nb = args->has_b ? args->timeseries_b.size() : na;
  const size_t m = args->window;
  // the caller pre-allocates the profiles (InitProfileMemory)
  const bool matrix = t == PROFILE_TYPE_MATRIX_SUMMARY;
  const size_t need_a =
      matrix ? (size_t)(args->matrix_height * args->matrix_width) : na - m + 1;
  if (args->profile_a.type != t || profile_length(args->profile_a, t) < need_a) {
    throw SCAMPException("Error: profile_a is not allocated for this join");
  }
  const bool sep = !matrix && args->computing_rows && args->keep_rows_separate;
  if (sep && (args->profile_b.type != t ||
              profile_length(args->profile_b, t) < nb - m + 1)) {
    throw SCAMPException("Error: profile_b is not allocated for this join");
  }

  scamp_hip_params p{};
  p.window = (int64_t)m;
  p.profile_type = t;
  p.precision = args->precision_type;
  p.has_b = args->has_b;
  p.computing_rows = args->computing_rows;
  p.computing_columns = args->computing_columns;
  p.keep_rows_separate = args->keep_rows_separate;
  p.is_aligned = args->is_aligned;
  p.silent_mode = args->silent_mode;
  p.distributed_start_row = args->distributed_start_row;
  p.distributed_start_col = args->distributed_start_col;
  p.distance_threshold = args->distance_threshold;
  p.matrix_height = args->matrix_height;
  p.matrix_width = args->matrix_width;
  p.max_tile_size = (int64_t)args->max_tile_size;
  p.anchor_interval = args->anchor_interval;
  p.max_matches_per_column = args->max_matches_per_column;

  CallContext call(device);
  scamp_hip_ctx *ctx = call.get();
  int rc;
  if (t == PROFILE_TYPE_APPROX_ALL_NEIGHBORS) {
    const int64_t k = args->max_matches_per_column;
    if (k < 1 || k > 256)
      throw SCAMPException("Error: max_matches_per_column must be between 1 and 256");
    std::vector<scamp_hip_match> ma = queues_to_matches(args->profile_a, na - m + 1, k);
    std::vector<scamp_hip_match> mb;
    if (sep) mb = queues_to_matches(args->profile_b, nb - m + 1, k);
    rc = scamp_hip_join(ctx, args->timeseries_a.data(), na,
                        args->has_b ? args->timeseries_b.data() : nullptr,
                        args->has_b ? nb : 0, &p, ma.data(), sep ? mb.data() : nullptr);
    if (rc == SCAMP_HIP_OK) {
      matches_to_queues(ma, args->profile_a, na - m + 1, k);
      if (sep) matches_to_queues(mb, args->profile_b, nb - m + 1, k);
    }
  } else {
    rc = scamp_hip_join(ctx, args->timeseries_a.data(), na,
                        args->has_b ? args->timeseries_b.data() : nullptr,
                        args->has_b ? nb : 0, &p, profile_buffer(args->profile_a, t),
                        sep ? profile_buffer(args->profile_b, t) : nullptr);
  }
  if (rc != SCAMP_HIP_OK) {
    char buf[512];
    scamp_hip_last_error(buf, sizeof(buf));
    throw SCAMPException(std::string(buf));
  }
}

void do_SCAMP(SCAMPArgs *args) { do_SCAMP(args, std::vector<int>(), 0); }

}  // namespace SCAMP

// src/scamp_utils.cpp:226-245
bool InitProfileMemory(SCAMP::SCAMPArgs *args) {
  int64_t profile_a_size = (int64_t)args->timeseries_a.size() - (int64_t)args->window + 1;
  int64_t profile_b_size =
      args->has_b ? (int64_t)args->timeseries_b.size() - (int64_t)args->window + 1
                  : profile_a_size;
  if (profile_a_size <= 0 || (args->keep_rows_separate && profile_b_size <= 0)) {
    return false;
  }
  args->profile_a.data.clear();
  args->profile_a.Alloc(profile_a_size, args->matrix_height, args->matrix_width,
                        args->distance_threshold);
  if (args->keep_rows_separate) {
    args->profile_b.data.clear();
    args->profile_b.Alloc(profile_b_size, args->matrix_height,
                          args->matrix_width, args->distance_threshold);
  }
  return true;
}

// src/scamp_utils.cpp:127-141
double ConvertToEuclidean(double val, int window) {
  if (val < -1) return NAN;
  return std::sqrt(std::max(2.0 * window * (1.0 - val), 0.0));
}

double CleanupPearson(double val) {
  if (val < -1) return NAN;
  return val;
}

// ---------------------------------------------------------------------------
// Text formats of the command-line tool (src/scamp_utils.cpp:1-224).

namespace {

void write_value(std::ofstream &out, double v) {
  out << std::setprecision(10) << v;
}

}  // namespace

bool readFile(const std::string &filename, std::vector<double> &v) {
  std::ifstream f(filename);
  if (f.fail()) {
    std::cout << "Unable to open" << filename
              << "for reading, please make sure it exists" << std::endl;
    return false;
  }
  // Token by token (blank lines are skipped); as in the reference reader
  // (src/scamp_utils.cpp:28-74) a value is kept only when some character
  // follows it, so a last value without a line end is dropped.
  std::string tok;
  while (f >> tok) {
    double d;
    try {
      d = std::stod(tok);
    } catch (const std::invalid_argument &) {
      std::cout << tok[0] << std::endl
                << "FATAL ERROR: invalid argument: Could not parse line number "
                << v.size() + 1 << " from input file.\n";
      std::exit(1);
    } catch (const std::out_of_range &) {
      std::cout << tok[0] << std::endl
                << "FATAL ERROR: out of range: Could not parse line number "
                << v.size() + 1 << " from input file.\n";
      std::exit(1);
    }
    if (f.peek() == EOF) break;
    v.push_back(d);
  }
  return true;
}

std::vector<int> ParseIntList(const std::string &s) {
  std::vector<int> out;
  if (s.empty()) return out;
  std::stringstream ss(s);
  std::string part;
  while (std::getline(ss, part, ',')) out.push_back(std::stoi(part));
  return out;
}

SCAMP::SCAMPPrecisionType GetPrecisionType(bool ultrap, bool doublep,
                                           bool mixedp, bool singlep) {
  if (ultrap) return SCAMP::PRECISION_ULTRA;
  if (doublep) return SCAMP::PRECISION_DOUBLE;
  if (mixedp) return SCAMP::PRECISION_MIXED;
  if (singlep) return SCAMP::PRECISION_SINGLE;
  return SCAMP::PRECISION_INVALID;
}

SCAMP::SCAMPProfileType ParseProfileType(const std::string &s) {
  if (s == "1NN_INDEX") return SCAMP::PROFILE_TYPE_1NN_INDEX;
  if (s == "SUM_THRESH") return SCAMP::PROFILE_TYPE_SUM_THRESH;
  if (s == "1NN") return SCAMP::PROFILE_TYPE_1NN;
  if (s == "ALL_NEIGHBORS") return SCAMP::PROFILE_TYPE_APPROX_ALL_NEIGHBORS;
  if (s == "MATRIX_SUMMARY") return SCAMP::PROFILE_TYPE_MATRIX_SUMMARY;
  return SCAMP::PROFILE_TYPE_INVALID;
}

bool WriteProfileToFile(const std::string &mp, const std::string &mpi,
                        SCAMP::Profile &p, bool output_pearson, int window,
                        int matrix_width, int matrix_height) {
  (void)matrix_height;
  auto convert = [&](double corr) {
    return output_pearson ? CleanupPearson(corr) : ConvertToEuclidean(corr, window);
  };
  switch (p.type) {
    case SCAMP::PROFILE_TYPE_1NN_INDEX: {
      std::ofstream out(mp), iout(mpi);
      for (const uint64_t u : p.data[0].uint64_value) {
        SCAMP::mp_entry e;
        e.ulong = u;
        write_value(out, convert(e.floats[0]));
        out << std::endl;
        iout << (e.floats[0] < -1 ? -1 : (int)e.ints[1]) << std::endl;
      }
      return true;
    }
    case SCAMP::PROFILE_TYPE_1NN: {
      std::ofstream out(mp);
      for (const float f : p.data[0].float_value) {
        write_value(out, convert(f));
        out << std::endl;
      }
      return true;
    }
    case SCAMP::PROFILE_TYPE_SUM_THRESH: {
      std::ofstream out(mp);
      for (const double d : p.data[0].double_value) {
        write_value(out, d);
        out << std::endl;
      }
      return true;
    }
    case SCAMP::PROFILE_TYPE_APPROX_ALL_NEIGHBORS: {
      // "col row value" per match, each subsequence's matches best first
      std::ofstream out(mp);
      for (auto q : p.data[0].match_value) {
        std::vector<SCAMP::SCAMPmatch> v;
        while (!q.empty()) {
          v.push_back(q.top());
          q.pop();
        }
        for (auto it = v.rbegin(); it != v.rend(); ++it) {
          out << it->col << " " << it->row << " ";
          write_value(out, convert(it->corr));
          out << std::endl;
        }
      }
      return true;
    }
    case SCAMP::PROFILE_TYPE_MATRIX_SUMMARY: {
      // rows of matrix_width values, each followed by a space
      std::ofstream out(mp);
      const std::vector<float> &m = p.data[0].float_value;
      for (size_t i = 0; i < m.size(); ++i) {
        if (i > 0 && i % (size_t)matrix_width == 0) out << std::endl;
        write_value(out, convert(m[i]));
        out << " ";
      }
      out << std::endl;
      return true;
    }
    default:
      return false;
  }
}
