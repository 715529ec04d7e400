// SCAMP's command-line tool on this engine: the same flags, checks, messages
// and output files as the reference's main (src/main.cpp:21-243), with the
// join run by SCAMP::do_SCAMP on the GPU.
//
//   SCAMP --window=100 --input_a_file_name=ts.txt [--input_b_file_name=...]
//         [--profile_type=1NN_INDEX|1NN|SUM_THRESH|MATRIX_SUMMARY]
//         [--double_precision|--mixed_precision|--single_precision|
//          --ultra_precision] [--output_pearson] [--keep_rows] [--aligned]
//         [--global_row=N --global_col=N] [--threshold=x]
//         [--reduced_height=H --reduced_width=W] [--gpus=0]
//         [--output_a_file_name=...] [--output_a_index_file_name=...]
//         [--output_b_file_name=...] [--output_b_index_file_name=...]
//         [--anchor_interval=R]
//
// Flags follow gflags' syntax: --name=value, --name value, -name=value, and
// for booleans --name, --noname, --name=true|false.  An unknown flag is an
// error.  --anchor_interval (rows between FP64 re-anchors of the mixed form)
// is this engine's own; --no_gpu is refused (there is no CPU path) and
// --num_cpu_workers is accepted and unused.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "scamp.hpp"

namespace {

struct Flag {
  enum Kind { kBool, kInt, kDouble, kString } kind;
  bool b = false;
  int64_t i = 0;
  double d = 0;
  std::string s;
};

std::map<std::string, Flag> make_flags() {
  std::map<std::string, Flag> f;
  auto B = [&](const char *n, bool v) { Flag x{Flag::kBool}; x.b = v; f[n] = x; };
  auto I = [&](const char *n, int64_t v) { Flag x{Flag::kInt}; x.i = v; f[n] = x; };
  auto D = [&](const char *n, double v) { Flag x{Flag::kDouble}; x.d = v; f[n] = x; };
  auto S = [&](const char *n, const char *v) { Flag x{Flag::kString}; x.s = v; f[n] = x; };
  I("max_matches_per_column", 5);
  I("reduced_height", 50);
  I("reduced_width", 50);
  B("print_debug_info", false);
  I("num_cpu_workers", 0);
  B("output_pearson", false);
  B("no_gpu", false);
  I("max_tile_size", 1 << 17);
  I("window", -1);
  D("threshold", 0);
  S("profile_type", "1NN_INDEX");
  B("ultra_precision", false);
  B("double_precision", false);
  B("mixed_precision", false);
  B("single_precision", false);
  B("keep_rows", false);
  B("aligned", false);
  I("global_row", -1);
  I("global_col", -1);
  S("input_a_file_name", "");
  S("input_b_file_name", "");
  S("output_a_file_name", "mp_columns_out");
  S("output_a_index_file_name", "mp_columns_out_index");
  S("output_b_file_name", "mp_rows_out");
  S("output_b_index_file_name", "mp_rows_out_index");
  S("gpus", "");
  I("anchor_interval", 0);
  return f;
}

[[noreturn]] void flag_error(const std::string &msg) {
  std::cerr << "ERROR: " << msg << std::endl;
  std::exit(1);
}

bool parse_bool(const std::string &v, const std::string &name) {
  if (v == "true" || v == "1" || v == "yes" || v == "t" || v == "y") return true;
  if (v == "false" || v == "0" || v == "no" || v == "f" || v == "n") return false;
  flag_error("illegal value '" + v + "' specified for bool flag '" + name + "'");
}

void set_value(Flag &f, const std::string &name, const std::string &v) {
  try {
    size_t used = 0;
    switch (f.kind) {
      case Flag::kBool:
        f.b = parse_bool(v, name);
        return;
      case Flag::kInt:
        f.i = std::stoll(v, &used);
        break;
      case Flag::kDouble:
        f.d = std::stod(v, &used);
        break;
      case Flag::kString:
        f.s = v;
        return;
    }
    if (used != v.size()) throw std::invalid_argument(v);
  } catch (const std::exception &) {
    flag_error("illegal value '" + v + "' specified for flag '" + name + "'");
  }
}

void parse_flags(int argc, char **argv, std::map<std::string, Flag> &flags) {
  for (int k = 1; k < argc; ++k) {
    std::string arg = argv[k];
    if (arg.size() < 2 || arg[0] != '-') continue;  // positional: ignored
    if (arg == "--") break;
    arg = arg.substr(arg[1] == '-' ? 2 : 1);
    std::string name = arg, value;
    const size_t eq = arg.find('=');
    const bool has_value = eq != std::string::npos;
    if (has_value) {
      name = arg.substr(0, eq);
      value = arg.substr(eq + 1);
    }
    auto it = flags.find(name);
    if (it == flags.end() && !has_value && name.rfind("no", 0) == 0) {
      auto jt = flags.find(name.substr(2));
      if (jt != flags.end() && jt->second.kind == Flag::kBool) {
        jt->second.b = false;
        continue;
      }
    }
    if (it == flags.end()) flag_error("unknown command line flag '" + name + "'");
    Flag &f = it->second;
    if (!has_value) {
      if (f.kind == Flag::kBool) {
        f.b = true;
        continue;
      }
      if (k + 1 >= argc) flag_error("flag '" + name + "' is missing its argument");
      value = argv[++k];
    }
    set_value(f, name, value);
  }
}

}  // namespace

int main(int argc, char **argv) {
  std::map<std::string, Flag> F = make_flags();
  parse_flags(argc, argv, F);
  auto B = [&](const char *n) { return F.at(n).b; };
  auto I = [&](const char *n) { return F.at(n).i; };
  auto S = [&](const char *n) { return F.at(n).s; };

  bool ultra = B("ultra_precision"), dbl = B("double_precision"),
       mixed = B("mixed_precision"), single = B("single_precision");
  if (!ultra && !dbl && !mixed && !single) dbl = true;
  if ((int)ultra + (int)dbl + (int)mixed + (int)single != 1) {
    printf("Error: only one precision flag can be enabled at a time\n");
    return 1;
  }
  const int64_t window = I("window");
  if (window < 3) {
    printf(
        "Error: Subsequence length must be at least 3, please use "
        "--window=<window_size> to specify your subsequence length.\n");
    return 1;
  }
  if (I("max_tile_size") < 1024) {
    printf("Error: max tile size must be at least 1024\n");
    return 1;
  }
  if (I("max_tile_size") / 2 < window) {
    printf(
        "Error: Tile length and width must be at least 2x larger than the "
        "window size. Please set a larger --max_tile_size=<max_tile_size>\n");
    return 1;
  }
  std::vector<int> devices;
  try {
    devices = ParseIntList(S("gpus"));
  } catch (const std::exception &) {
    printf("Error: could not parse --gpus=%s\n", S("gpus").c_str());
    return 1;
  }
  if (B("no_gpu")) {
    printf("Error: this engine computes on the GPU only (--no_gpu is not supported)\n");
    return 1;
  }
  if (S("input_a_file_name").empty()) {
    printf(
        "Error: primary input filename must be specified using "
        "--input_a_file_name");
    return 1;
  }
  const bool self_join = S("input_b_file_name").empty();
  const bool computing_cols = true;
  const bool computing_rows = self_join ? true : B("keep_rows");

  const SCAMP::SCAMPPrecisionType precision = GetPrecisionType(ultra, dbl, mixed, single);
  const SCAMP::SCAMPProfileType profile_type = ParseProfileType(S("profile_type"));

  std::vector<double> ta, tb;
  if (!readFile(S("input_a_file_name"), ta)) return 1;
  if (!self_join && !readFile(S("input_b_file_name"), tb)) return 1;

  const int64_t n_x = (int64_t)ta.size() - window + 1;
  const int64_t n_y = self_join ? n_x : (int64_t)tb.size() - window + 1;
  if (n_x < 1 || n_y < 1) {
    printf("Error: window size must be smaller than the timeseries length\n");
    return 1;
  }

  SCAMP::SCAMPArgs args;
  args.window = window;
  args.max_tile_size = I("max_tile_size");
  args.has_b = !self_join;
  args.distributed_start_row = I("global_row");
  args.distributed_start_col = I("global_col");
  args.distance_threshold = F.at("threshold").d;
  args.computing_columns = computing_cols;
  args.computing_rows = computing_rows;
  args.profile_a.type = profile_type;
  args.profile_b.type = profile_type;
  args.precision_type = precision;
  args.profile_type = profile_type;
  args.keep_rows_separate = B("keep_rows");
  args.is_aligned = B("aligned");
  args.timeseries_a = std::move(ta);
  args.timeseries_b = std::move(tb);
  args.silent_mode = !B("print_debug_info");
  args.max_matches_per_column = I("max_matches_per_column");
  args.matrix_height = I("reduced_height");
  args.matrix_width = I("reduced_width");
  args.anchor_interval = I("anchor_interval");
  if (B("print_debug_info")) printf("Starting SCAMP\n");
  try {
    InitProfileMemory(&args);
    if (devices.empty())
      SCAMP::do_SCAMP(&args);
    else
      SCAMP::do_SCAMP(&args, devices, (int)I("num_cpu_workers"));
  } catch (const SCAMPException &e) {
    std::cout << e.what() << "\n";
    return 1;
  }
  if (B("print_debug_info")) printf("Now writing result to files\n");
  WriteProfileToFile(S("output_a_file_name"), S("output_a_index_file_name"),
                     args.profile_a, B("output_pearson"), (int)window,
                     (int)I("reduced_width"), (int)I("reduced_height"));
  if (B("keep_rows")) {
    WriteProfileToFile(S("output_b_file_name"), S("output_b_index_file_name"),
                       args.profile_b, B("output_pearson"), (int)window,
                       (int)I("reduced_width"), (int)I("reduced_height"));
  }
  if (B("print_debug_info")) printf("Done\n");
  return 0;
}
