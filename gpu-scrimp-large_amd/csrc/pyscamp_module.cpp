// pyscamp: the Python surface of SCAMP (python/SCAMP_python.cpp:396-548) on
// the MI355X engine.  Same function names, positional arguments, keyword
// whitelists per profile type, defaults and output dtypes:
//   selfjoin(a, m, **kw)          -> (float32[n_sub] distance, int32[n_sub] index)
//   abjoin(a, b, m, **kw)         -> (float32[n_sub(a)], int32[n_sub(a)] into b)
//   selfjoin_sum / abjoin_sum     -> float64[n_sub(a)]
//   selfjoin_matrix / abjoin_matrix -> float32[mheight, mwidth]
//   selfjoin_knn(a, m, k) / abjoin_knn(a, b, m, k)
//                                 -> [(col, row, distance), ...]: for each
//                                    subsequence col of a, its k best matches
//                                    above the threshold, worst first
//   gpu_supported()               -> bool
// Keywords: verbose, precision ('single'|'mixed'|'double'|'ultra'), pearson,
// gpus, threads, plus threshold (sum/knn/matrix) and mheight/mwidth (matrix);
// anything else raises ValueError, as in the reference.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cmath>
#include <set>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "scamp.hpp"

namespace py = pybind11;

namespace {

using Series = py::array_t<double, py::array::c_style | py::array::forcecast>;

std::vector<double> to_vector(const Series &a) {
  auto r = a.unchecked<1>();
  std::vector<double> v(r.shape(0));
  for (py::ssize_t i = 0; i < r.shape(0); ++i) v[i] = r(i);
  return v;
}

// GetDefaultSCAMPArgs (python/SCAMP_python.cpp:67-89)
SCAMP::SCAMPArgs default_args(SCAMP::SCAMPProfileType type) {
  SCAMP::SCAMPArgs args;
  args.has_b = false;
  args.max_tile_size = 128000;
  args.distributed_start_row = -1;
  args.distributed_start_col = -1;
  args.distance_threshold = 0;
  args.precision_type = SCAMP::PRECISION_DOUBLE;
  args.profile_type = type;
  args.computing_rows = true;
  args.computing_columns = true;
  args.keep_rows_separate = false;
  args.is_aligned = false;
  args.silent_mode = true;
  args.max_matches_per_column = 5;
  args.matrix_height = 50;
  args.matrix_width = 50;
  args.profile_a.type = type;
  args.profile_b.type = type;
  return args;
}

// KeyIsOkForProfileType (python/SCAMP_python.cpp:91-114)
bool key_ok(const std::string &key, SCAMP::SCAMPProfileType type) {
  static const std::set<std::string> nn_index = {"verbose", "precision",
                                                 "pearson", "gpus", "threads"};
  static const std::set<std::string> sum_thresh = {
      "verbose", "precision", "pearson", "gpus", "threads", "threshold"};
  static const std::set<std::string> knn = {
      "verbose", "precision", "pearson", "gpus", "threads", "threshold"};
  static const std::set<std::string> matrix = {
      "verbose", "precision", "pearson", "gpus",
      "threads", "threshold", "mheight", "mwidth"};
  switch (type) {
    case SCAMP::PROFILE_TYPE_1NN_INDEX: return nn_index.count(key) == 1;
    case SCAMP::PROFILE_TYPE_SUM_THRESH: return sum_thresh.count(key) == 1;
    case SCAMP::PROFILE_TYPE_APPROX_ALL_NEIGHBORS: return knn.count(key) == 1;
    case SCAMP::PROFILE_TYPE_MATRIX_SUMMARY: return matrix.count(key) == 1;
    default: return false;
  }
}

// get_args_based_on_kwargs (python/SCAMP_python.cpp:116-177)
void apply_kwargs(SCAMP::SCAMPArgs *args, const py::kwargs &kwargs,
                  bool &pearson, std::vector<int> &gpus, int &threads) {
  for (auto item : kwargs) {
    const std::string key = py::str(item.first);
    if (!key_ok(key, args->profile_type))
      throw std::invalid_argument(
          "Invalid keyword argument specified unknown argument: " + key);
    if (key == "threshold") {
      args->distance_threshold = item.second.cast<double>();
      if (args->distance_threshold > 1 || args->distance_threshold < -1)
        throw std::invalid_argument(
            "Invalid threshold specified: value must be between -1 and 1");
    } else if (key == "verbose") {
      args->silent_mode = !item.second.cast<bool>();
    } else if (key == "mheight") {
      args->matrix_height = item.second.cast<int>();
      if (args->matrix_height <= 0)
        throw std::invalid_argument(
            "Invalid matrix height specified: value must be greater than 0");
    } else if (key == "mwidth") {
      args->matrix_width = item.second.cast<int>();
      if (args->matrix_width <= 0)
        throw std::invalid_argument(
            "Invalid matrix width specified: value must be greater than 0");
    } else if (key == "precision") {
      const std::string p = item.second.cast<std::string>();
      if (p == "single") {
        args->precision_type = SCAMP::PRECISION_SINGLE;
      } else if (p == "mixed") {
        args->precision_type = SCAMP::PRECISION_MIXED;
      } else if (p == "double") {
        args->precision_type = SCAMP::PRECISION_DOUBLE;
      } else if (p == "ultra") {
        args->precision_type = SCAMP::PRECISION_ULTRA;
      } else {
        throw std::invalid_argument(
            "Invalid precision type specified: valid options are single, "
            "mixed, double, ultra");
      }
    } else if (key == "pearson") {
      pearson = item.second.cast<bool>();
    } else if (key == "gpus") {
      gpus = item.second.cast<std::vector<int>>();
    } else if (key == "threads") {
      threads = item.second.cast<int>();
      if (threads < 0)
        throw std::invalid_argument(
            "Invalid number of cpu worker threads specified, must be greater "
            "than or equal to 0.");
    }
  }
}

bool run(SCAMP::SCAMPArgs *args, const py::kwargs &kwargs) {
  std::vector<int> gpus;
  int threads = 0;
  bool pearson = false;
  if (kwargs) apply_kwargs(args, kwargs, pearson, gpus, threads);
  InitProfileMemory(args);
  {
    py::gil_scoped_release release;
    SCAMP::do_SCAMP(args, gpus, threads);
  }
  return pearson;
}

py::tuple split_1nn_index(const std::vector<uint64_t> &p, bool pearson,
                          int window) {
  py::array_t<float> nn(p.size());
  py::array_t<int> idx(p.size());
  auto n = nn.mutable_unchecked<1>();
  auto x = idx.mutable_unchecked<1>();
  for (size_t i = 0; i < p.size(); ++i) {
    SCAMP::mp_entry e;
    e.ulong = p[i];
    n(i) = pearson ? CleanupPearson(e.floats[0])
                   : ConvertToEuclidean(e.floats[0], window);
    x(i) = (int)e.ints[1];
  }
  return py::make_tuple(nn, idx);
}

template <typename T>
py::array_t<T> to_array(const std::vector<T> &v, bool pearson = true,
                        int window = 0) {
  py::array_t<T> out(v.size());
  auto o = out.template mutable_unchecked<1>();
  for (size_t i = 0; i < v.size(); ++i)
    o(i) = pearson ? CleanupPearson(v[i]) : ConvertToEuclidean(v[i], window);
  return out;
}

py::tuple selfjoin(const Series &a, int m, const py::kwargs &kw) {
  auto args = default_args(SCAMP::PROFILE_TYPE_1NN_INDEX);
  args.timeseries_a = to_vector(a);
  args.timeseries_b = args.timeseries_a;
  args.window = m;
  bool pearson = run(&args, kw);
  return split_1nn_index(args.profile_a.data[0].uint64_value, pearson, m);
}

py::tuple abjoin(const Series &a, const Series &b, int m, const py::kwargs &kw) {
  auto args = default_args(SCAMP::PROFILE_TYPE_1NN_INDEX);
  args.timeseries_a = to_vector(a);
  args.timeseries_b = to_vector(b);
  args.window = m;
  args.has_b = true;
  args.computing_rows = false;
  bool pearson = run(&args, kw);
  return split_1nn_index(args.profile_a.data[0].uint64_value, pearson, m);
}

py::array_t<double> selfjoin_sum(const Series &a, int m, const py::kwargs &kw) {
  auto args = default_args(SCAMP::PROFILE_TYPE_SUM_THRESH);
  args.timeseries_a = to_vector(a);
  args.timeseries_b = args.timeseries_a;
  args.window = m;
  run(&args, kw);
  return to_array<double>(args.profile_a.data[0].double_value);
}

py::array_t<double> abjoin_sum(const Series &a, const Series &b, int m,
                               const py::kwargs &kw) {
  auto args = default_args(SCAMP::PROFILE_TYPE_SUM_THRESH);
  args.timeseries_a = to_vector(a);
  args.timeseries_b = to_vector(b);
  args.window = m;
  args.has_b = true;
  args.computing_rows = false;
  run(&args, kw);
  return to_array<double>(args.profile_a.data[0].double_value);
}

py::array_t<float> matrix_out(SCAMP::SCAMPArgs &args, bool pearson, int m) {
  auto arr = to_array<float>(args.profile_a.data[0].float_value, pearson, m);
  arr.resize({(py::ssize_t)args.matrix_height, (py::ssize_t)args.matrix_width});
  return arr;
}

py::array_t<float> selfjoin_matrix(const Series &a, int m, const py::kwargs &kw) {
  auto args = default_args(SCAMP::PROFILE_TYPE_MATRIX_SUMMARY);
  args.timeseries_a = to_vector(a);
  args.timeseries_b = args.timeseries_a;
  args.window = m;
  bool pearson = run(&args, kw);
  return matrix_out(args, pearson, m);
}

py::array_t<float> abjoin_matrix(const Series &a, const Series &b, int m,
                                 const py::kwargs &kw) {
  auto args = default_args(SCAMP::PROFILE_TYPE_MATRIX_SUMMARY);
  args.timeseries_a = to_vector(a);
  args.timeseries_b = to_vector(b);
  args.window = m;
  args.has_b = true;
  args.computing_rows = false;
  bool pearson = run(&args, kw);
  return matrix_out(args, pearson, m);
}

// SplitProfileKNN (python/SCAMP_python.cpp:31-50): each subsequence's kept
// matches in heap order (worst first), as (col, row, distance) tuples.
std::vector<std::tuple<int64_t, int64_t, float>> split_knn(SCAMP::SCAMPArgs &args,
                                                           bool pearson) {
  std::vector<std::tuple<int64_t, int64_t, float>> out;
  for (auto &q : args.profile_a.data[0].match_value) {
    while (!q.empty()) {
      const SCAMP::SCAMPmatch &e = q.top();
      const float d = pearson ? (float)CleanupPearson(e.corr)
                              : (float)ConvertToEuclidean(e.corr, (int)args.window);
      out.emplace_back((int64_t)e.col, (int64_t)e.row, d);
      q.pop();
    }
  }
  return out;
}

std::vector<std::tuple<int64_t, int64_t, float>> selfjoin_knn(const Series &a, int m,
                                                              int k,
                                                              const py::kwargs &kw) {
  auto args = default_args(SCAMP::PROFILE_TYPE_APPROX_ALL_NEIGHBORS);
  args.timeseries_a = to_vector(a);
  args.timeseries_b = args.timeseries_a;
  args.window = m;
  args.max_matches_per_column = k;
  bool pearson = run(&args, kw);
  return split_knn(args, pearson);
}

std::vector<std::tuple<int64_t, int64_t, float>> abjoin_knn(const Series &a,
                                                            const Series &b, int m,
                                                            int k,
                                                            const py::kwargs &kw) {
  auto args = default_args(SCAMP::PROFILE_TYPE_APPROX_ALL_NEIGHBORS);
  args.timeseries_a = to_vector(a);
  args.timeseries_b = to_vector(b);
  args.window = m;
  args.has_b = true;
  args.computing_rows = false;
  args.max_matches_per_column = k;
  bool pearson = run(&args, kw);
  return split_knn(args, pearson);
}

}  // namespace

PYBIND11_MODULE(pyscamp, m) {
  m.doc() = "pyscamp: SCAMP's Python interface on the MI355X matrix-profile engine";
  py::register_exception<SCAMPException>(m, "SCAMPException", PyExc_RuntimeError);
  m.def("gpu_supported", []() { return true; },
        "Returns whether or not the module was compiled with GPU support");
  m.def("selfjoin", &selfjoin, py::arg("a"), py::arg("m"),
        "Matrix profile (distance, index) of time series a");
  m.def("abjoin", &abjoin, py::arg("a"), py::arg("b"), py::arg("m"),
        "For each subsequence of a, its nearest neighbour in b");
  m.def("selfjoin_sum", &selfjoin_sum, py::arg("a"), py::arg("m"),
        "Sum of correlations above the threshold for each subsequence of a");
  m.def("abjoin_sum", &abjoin_sum, py::arg("a"), py::arg("b"), py::arg("m"),
        "Sum of correlations above the threshold to subsequences of b");
  m.def("selfjoin_matrix", &selfjoin_matrix, py::arg("a"), py::arg("m"),
        "Max-pooled distance matrix of shape (mheight, mwidth)");
  m.def("abjoin_matrix", &abjoin_matrix, py::arg("a"), py::arg("b"),
        py::arg("m"), "Max-pooled AB distance matrix of shape (mheight, mwidth)");
  m.def("selfjoin_knn", &selfjoin_knn, py::arg("a"), py::arg("m"), py::arg("k"),
        "For each subsequence of a, its k best matches in a (list of (col, row, "
        "distance))");
  m.def("abjoin_knn", &abjoin_knn, py::arg("a"), py::arg("b"), py::arg("m"),
        py::arg("k"),
        "For each subsequence of a, its k best matches in b (list of (col, row, "
        "distance))");
  m.attr("__version__") = "mi355x";
}
