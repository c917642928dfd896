// `ngap` command-line tool: a drop-in for the reference binary
// (code/src/ngap/main.cu:16-137) on the MI355X engine.
//
// Same flags and defaults (code/include/commons/common_func.h:50-115,
// code/src/ngap/ngap_option.h:12-68), same --algorithm names, same input
// slicing / duplication / chunking, and the stdout lines downstream scripts
// parse: "Results number:", "Unique results number:", "Validation PASS!" /
// "Validation FAILED!", "Result(N):" / "Reference result(N):" with the first
// 40 reports as 0x%lx, "ngap elapsed time: S seconds, throughput = X MB/s"
// (or WRONGTHRPUT(X)), "FINISHED!" (code/src/ngap/ngap.cu:1827-1990,
// 2270-2344).  Added flags: --schedule=plain|batched, --window=N,
// --write-reports=true (writes the sorted reports to --report-filename; the
// reference accepts the flag but never writes the file).
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "../../include/ngap_amd.h"
#include "frontend.h"

namespace ngapamd {
std::vector<uint64_t> cpu_reference(const std::vector<GroupGraph> &gs, const uint8_t *input,
                                    uint64_t chunk_len, uint32_t num_seg, bool is_dup);
}
using namespace ngapamd;

#define RESET "\033[0m"
#define BOLDRED "\033[1m\033[31m"
#define BOLDGREEN "\033[1m\033[32m"
#define BOLDBLUE "\033[1m\033[34m"

namespace {

struct Options {
  // common_gpunfa_options
  int block_size = 256;
  std::string algorithm = "graph";
  std::string report_filename = "report.txt";
  int max_nfa_size = -1;
  int duplicate_input_stream = 1;
  int input_start_pos = 0;
  int input_len = -1;
  unsigned long long output_capacity = 54619400ull;
  bool report_off = false;
  bool show_help = false;
  int split_chunk_size = -1;
  int only_exec_cc = -1;
  std::string input_filename, nfa_filename, select_cc_by_state, app_name;
  int padding = 256;
  long long quick_validation = -1;
  int unique_frequency = 0;
  bool use_uvm = false, quit_degree = true, remove_degree = false;
  // ngap_option
  unsigned data_buffer_fetch_size = 128;
  int add_aas_start = 0, add_aas_interval = 1;
  bool unique = false, validation = true;
  int active_threshold = 20;
  bool use_soa = false;
  int precompute_cutoff = -1, precompute_depth = 0;
  bool motivate_worklist_length = false;
  int num_state_per_group = 256;
  int group_num = 10;
  bool compress_prec_table = true, tuning = false, pc_use_uvm = false, adaptive_aas = false,
       try_adaptive_aas = false;
  // added
  std::string schedule = "batched";
  int window = 0;
  bool write_reports = false;
  unsigned spill_capacity = 0;
};

bool parse_bool(const std::string &v, bool &out) {
  std::string s = v;
  std::transform(s.begin(), s.end(), s.begin(), ::tolower);
  if (s == "y" || s == "1" || s == "true" || s == "yes" || s == "on") return out = true, true;
  if (s == "n" || s == "0" || s == "false" || s == "no" || s == "off") return out = false, true;
  return false;
}

template <class T>
bool parse_num(const std::string &v, T &out) {
  char *end = nullptr;
  long long x = strtoll(v.c_str(), &end, 10);
  if (!end || *end || v.empty()) return false;
  out = (T)x;
  return true;
}

bool parse_args(int argc, char **argv, Options &o, std::string &err) {
  std::map<std::string, std::function<bool(const std::string &)>> opt;
  auto I = [&](int &x) { return [&x](const std::string &v) { return parse_num(v, x); }; };
  auto B = [&](bool &x) { return [&x](const std::string &v) { return parse_bool(v, x); }; };
  auto S = [&](std::string &x) { return [&x](const std::string &v) { x = v; return true; }; };
  opt["--algorithm"] = opt["-g"] = S(o.algorithm);
  opt["--block-size"] = I(o.block_size);
  opt["--max-nfa-size"] = I(o.max_nfa_size);
  opt["--report-off"] = B(o.report_off);
  opt["--report-filename"] = S(o.report_filename);
  opt["--result-capacity"] = [&](const std::string &v) { return parse_num(v, o.output_capacity); };
  opt["--input-start-pos"] = I(o.input_start_pos);
  opt["--input-len"] = I(o.input_len);
  opt["--input"] = opt["-i"] = S(o.input_filename);
  opt["--automata"] = opt["-a"] = S(o.nfa_filename);
  opt["--split-entire-inputstream-to-chunk-size"] = I(o.split_chunk_size);
  opt["--only-exec-cc-with-state-id"] = S(o.select_cc_by_state);
  opt["--only-exec-ccid"] = I(o.only_exec_cc);
  opt["--duplicate-input-stream"] = I(o.duplicate_input_stream);
  opt["--quick-validation"] = [&](const std::string &v) { return parse_num(v, o.quick_validation); };
  opt["--unique-frequency"] = I(o.unique_frequency);
  opt["--padding"] = I(o.padding);
  opt["--app-name"] = S(o.app_name);
  opt["--use-uvm"] = B(o.use_uvm);
  opt["--quit-degree"] = B(o.quit_degree);
  opt["--remove-degree"] = B(o.remove_degree);
  opt["--add-aan-start"] = I(o.add_aas_start);
  opt["--add-aas-interval"] = I(o.add_aas_interval);
  opt["--unique"] = B(o.unique);
  opt["--active-threshold"] = I(o.active_threshold);
  opt["--validation"] = B(o.validation);
  opt["--use-soa"] = B(o.use_soa);
  opt["--precompute-cutoff"] = I(o.precompute_cutoff);
  opt["--precompute-depth"] = I(o.precompute_depth);
  opt["--data-buffer-fetch-size"] = [&](const std::string &v) {
    return parse_num(v, o.data_buffer_fetch_size);
  };
  opt["--motivate-worklist-length"] = B(o.motivate_worklist_length);
  opt["--num-state-per-group"] = I(o.num_state_per_group);
  opt["--group-num"] = I(o.group_num);
  opt["--tuning"] = B(o.tuning);
  opt["--pc-use-uvm"] = B(o.pc_use_uvm);
  opt["--adaptive-aas"] = B(o.adaptive_aas);
  opt["--try-adaptive-aas"] = B(o.try_adaptive_aas);
  opt["--compress-prec-table"] = B(o.compress_prec_table);
  opt["--schedule"] = S(o.schedule);
  opt["--window"] = I(o.window);
  opt["--write-reports"] = B(o.write_reports);
  opt["--spill-capacity"] = [&](const std::string &v) { return parse_num(v, o.spill_capacity); };
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    if (a == "-h" || a == "-?" || a == "--help") {
      o.show_help = true;
      continue;
    }
    std::string key = a, val;
    bool has_val = false;
    size_t eq = a.find('=');
    if (a.size() > 1 && a[0] == '-' && eq != std::string::npos) {
      key = a.substr(0, eq);
      val = a.substr(eq + 1);
      has_val = true;
    }
    auto it = opt.find(key);
    if (it == opt.end()) {
      err = "Unrecognised token: " + a;
      return false;
    }
    if (!has_val) {
      if (i + 1 >= argc) {
        err = "Expected argument following " + key;
        return false;
      }
      val = argv[++i];
    }
    if (!it->second(val)) {
      err = "Unable to convert '" + val + "' for option " + key;
      return false;
    }
  }
  if (o.input_filename.empty() || o.nfa_filename.empty()) {
    err = "the input stream (-i) and the automata (-a) are required";
    return false;
  }
  return true;
}

void print_help() {
  std::cout << "usage: ngap -a <file.anml> -i <input> --algorithm=<name> [options]\n"
               "  algorithms: blockinggroups NAPgroups nonblockinggroups nonblockingr1groups\n"
               "              nonblockingr2groups nonblockingpcgroups nonblockingallgroups\n"
               "  options of the reference ngap binary are accepted with the same names and\n"
               "  defaults; added: --schedule=plain|batched --window=N --write-reports=true\n"
               "  --spill-capacity=N (worklist items per wave beyond its LDS ring)\n";
}

bool read_file(const std::string &fn, std::vector<uint8_t> &out) {
  std::ifstream f(fn, std::ios::binary);
  if (!f) return false;
  out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return true;
}

const std::map<std::string, int> kAlgorithms = {
    {"blockinggroups", NGAP_ALGO_BLOCKING},    {"NAPgroups", NGAP_ALGO_NAP},
    {"nonblockinggroups", NGAP_ALGO_NONBLOCKING}, {"nonblockingr1groups", NGAP_ALGO_R1},
    {"nonblockingr2groups", NGAP_ALGO_R2},     {"nonblockingpcgroups", NGAP_ALGO_PC},
    {"nonblockingallgroups", NGAP_ALGO_ALL}};

void print_vec(const std::vector<uint64_t> &v) {
  size_t n = std::min<size_t>(v.size(), 40);
  for (size_t i = 0; i < n; i++) printf("0x%lx, ", (unsigned long)v[i]);
  printf(v.size() > 40 ? "...\n" : "\n");
}

// automataValidation (code/src/ngap/ngap.cu:2270-2344)
bool validate(const std::vector<uint64_t> &res, const std::vector<uint64_t> &ref) {
  auto both = [&] {
    printf("Result(%zu): \n", res.size());
    print_vec(res);
    printf("Reference result(%zu): \n", ref.size());
    print_vec(ref);
    printf("\n");
  };
  bool ok = res == ref;
  if (!ok) {
    std::cout << BOLDRED << "Validation FAILED!" << RESET << "\n";
    std::cerr << BOLDRED << "Validation FAILED!" << RESET << "\n";
    size_t bound = std::min(res.size(), ref.size());
    for (size_t i = 0; i < bound; i++)
      if (res[i] != ref[i]) {
        printf("0x%lx != 0x%lx\n", (unsigned long)res[i], (unsigned long)ref[i]);
        break;
      }
    both();
    return false;
  }
  std::cout << BOLDGREEN << "Validation PASS!" << RESET << "\n";
  both();
  return true;
}

}  // namespace

int main(int argc, char **argv) {
  printf("Command: ");
  for (int i = 0; i < argc; i++) printf("%s ", argv[i]);
  printf("\n");
  Options o;
  std::string err;
  if (!parse_args(argc, argv, o, err)) {
    std::cerr << "Error in command line: " << err << std::endl;
    if (o.show_help) print_help();
    return 1;
  }
  if (o.show_help) print_help();

  std::vector<uint8_t> raw;
  std::cout << "read input stream from file = " << o.input_filename << std::endl;
  if (!read_file(o.input_filename, raw)) {
    std::cerr << "cannot read input stream " << o.input_filename << std::endl;
    return 1;
  }
  std::cout << "input_stream_size = " << raw.size() << std::endl;
  // slice only when both start and length are given (main.cu:49-52)
  if (o.input_start_pos != -1 && o.input_len != -1) {
    if (o.input_start_pos < 0 || o.input_len < 0 || (size_t)o.input_start_pos >= raw.size()) {
      std::cerr << "bad --input-start-pos / --input-len" << std::endl;
      return 1;
    }
    size_t len = o.input_len;
    if (o.input_start_pos + len > raw.size()) {
      len = raw.size() - o.input_start_pos;
      std::cout << "the input is shorter than the length specified, just slice to end"
                << std::endl;
    }
    raw = std::vector<uint8_t>(raw.begin() + o.input_start_pos,
                               raw.begin() + o.input_start_pos + len);
  }

  std::cout << "load automata from file = " << o.nfa_filename << std::endl;
  Nfa nfa;
  if (!load_anml_file(o.nfa_filename, nfa, err)) {
    std::cout << err << std::endl;
    return 1;
  }
  FrontendOptions fo;
  fo.group_num = o.group_num;
  fo.max_nfa_size = o.max_nfa_size;
  fo.quit_degree = o.quit_degree;
  fo.remove_degree = o.remove_degree;
  std::vector<GroupGraph> groups;
  FrontendStats fs;
  if (!build_groups(nfa, fo, groups, fs, err)) {
    std::cout << err << std::endl;
    return 1;
  }
  // print_nfa_info (code/src/commons/nfa_utils.cpp:946-986)
  std::cout << "NFA Info:\n    total_state_number = " << fs.total_states
            << "\n    start_state_number = " << fs.start_states
            << "\n    always_active_state_number = " << fs.always_active
            << "\n    report_state_number = " << fs.report_states
            << "\n    total_cc_number = " << fs.num_cc << "\n    cc_256_number = " << fs.cc_256
            << "\n    max_cc_size = " << fs.max_cc << "\n    average_cc_size = " << fs.avg_cc
            << std::endl;
  printf("ReadANML finish \n");
  printf("limit_out_degree_on_ccs: num_succeed %d, num_fail %d\n", fs.limit_succeed,
         fs.limit_fail);
  if (o.group_num > fo.group_num)
    printf("warning: cc_num = %d is less than group_num = %d, new group_num will be %d\n",
           fo.group_num, o.group_num, fo.group_num);
  o.group_num = fo.group_num;
  printf("grouped_nfas.size = %zu pl.num_seg=%d\n", groups.size(), 0);

  std::cout << "Input Stream Info:\n    input_start_pos = " << o.input_start_pos
            << "\n    input_length = " << o.input_len
            << "\n    split_entire_inputstream_to_chunk_size = " << o.split_chunk_size
            << "\n    dup_input_stream = " << o.duplicate_input_stream << std::endl;
  std::vector<uint8_t> ss;
  for (int i = 0; i < o.duplicate_input_stream; i++) ss.insert(ss.end(), raw.begin(), raw.end());
  if (o.split_chunk_size <= 0) {
    std::cerr << "--split-entire-inputstream-to-chunk-size must be positive" << std::endl;
    return 1;
  }
  uint32_t num_seg = (uint32_t)(ss.size() / o.split_chunk_size);
  if (num_seg == 0) {
    std::cerr << "input shorter than one chunk" << std::endl;
    return 1;
  }
  uint64_t chunk = (uint64_t)o.split_chunk_size;
  const uint64_t total = chunk * num_seg;
  bool is_dup = o.duplicate_input_stream > 1 &&
                (o.split_chunk_size == -1 || o.split_chunk_size == o.input_len);

  auto algo = kAlgorithms.find(o.algorithm);
  if (algo == kAlgorithms.end()) {
    std::cout << "not supported algoritm " << o.algorithm << std::endl;
    printf("FINISHED!\n");
    return 0;
  }
  std::cout << BOLDBLUE << "automata " << o.algorithm << " ::launch!" << RESET << "\n";

  auto t0 = std::chrono::high_resolution_clock::now();
  ngap_engine *eng = nullptr;
  if (ngap_create(0, &eng) != NGAP_OK) {
    std::cerr << ngap_last_error() << std::endl;
    return 1;
  }
  std::vector<ngap_group_desc> desc(groups.size());
  for (size_t g = 0; g < groups.size(); g++) {
    const GroupGraph &gg = groups[g];
    desc[g] = {gg.num_states, gg.row.data(), gg.col.data(), gg.sets.data(), gg.attrs.data(),
               (int32_t)gg.aas.size(), gg.aas.data(), (int32_t)gg.sod.size(), gg.sod.data()};
  }
  int depth = std::max(1, std::min(3, o.precompute_depth));
  if (ngap_load_groups(eng, desc.data(), (int32_t)desc.size(), depth) != NGAP_OK ||
      ngap_set_input(eng, ss.data(), chunk, num_seg) != NGAP_OK) {
    std::cerr << ngap_last_error() << std::endl;
    return 1;
  }
  auto t1 = std::chrono::high_resolution_clock::now();
  printf("prepare time = %f s\n",
         std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count() / 1e6);

  ngap_run_opts ro;
  ngap_default_opts(&ro);
  ro.algorithm = algo->second;
  ro.schedule = o.schedule == "plain" ? NGAP_SCHED_PLAIN : NGAP_SCHED_BATCHED;
  ro.precompute_depth = o.precompute_depth;
  ro.active_threshold = o.active_threshold;
  ro.unique = o.unique;
  ro.is_dup = is_dup;
  ro.report_off = o.report_off;
  ro.data_buffer_fetch_size = o.data_buffer_fetch_size;
  ro.add_aas_start = o.add_aas_start;
  ro.add_aas_interval = o.add_aas_interval;
  ro.result_capacity = o.output_capacity;
  ro.window = o.window;
  ro.spill_capacity = o.spill_capacity;
  // reference scheduling knobs this engine has no counterpart for: the
  // all-input roots of a step cover a wave's width of positions (one for
  // NAPgroups) and the worklist is the wave's LDS ring plus its spill area
  if (o.add_aas_start != 0 || o.add_aas_interval != 1 || o.data_buffer_fetch_size != 128 ||
      o.adaptive_aas)
    fprintf(stderr,
            "note: --add-aan-start, --add-aas-interval, --data-buffer-fetch-size and "
            "--adaptive-aas are accepted and have no effect on this engine\n");

  std::vector<uint64_t> ref;
  bool ref_done = false;
  auto one_run = [&](const ngap_run_opts &opts, bool &pass) -> double {
    ngap_run_stats st;
    int rc = ngap_run(eng, &opts, &st);
    double ms = st.kernel_ms;
    double throughput = (double)total / (ms * 1000.0);
    std::cout << "Results number: " << st.reports_raw << std::endl;
    pass = rc == NGAP_OK;
    if (rc != NGAP_OK) std::cout << BOLDRED << ngap_last_error() << RESET << "\n";
    if (pass && o.validation) {
      if (o.report_off) {
        unsigned long long dup = o.duplicate_input_stream > 1 ? o.duplicate_input_stream : 1;
        unsigned long long want = (unsigned long long)o.quick_validation * dup;
        pass = o.quick_validation >= 0 && want <= st.reports_raw;
        std::cout << (pass ? BOLDGREEN : BOLDRED)
                  << (pass ? (want == st.reports_raw ? "Quick Validation PASS! (report off, perfect)"
                                                     : "Quick Validation PASS! (report off, not perfect)")
                           : "Quick Validation FAILED! (report off)")
                  << RESET << "\n";
      } else {
        uint64_t n = 0;
        ngap_get_reports(eng, nullptr, 0, &n);
        std::vector<uint64_t> res(n);
        ngap_get_reports(eng, res.data(), n, &n);
        std::cout << "Unique results number: " << res.size() << std::endl;
        printf("validation start.\n");
        if (o.quick_validation >= 0) {
          unsigned long long dup = o.duplicate_input_stream > 1 ? o.duplicate_input_stream : 1;
          pass = (unsigned long long)o.quick_validation * dup == res.size();
          std::cout << (pass ? BOLDGREEN : BOLDRED)
                    << (pass ? "Quick Validation PASS!" : "Quick Validation FAILED!") << RESET
                    << "\n";
        } else {
          if (!ref_done) {
            auto r0 = std::chrono::high_resolution_clock::now();
            ref = cpu_reference(groups, ss.data(), chunk, num_seg, is_dup);
            auto r1 = std::chrono::high_resolution_clock::now();
            auto ms_ref = std::chrono::duration_cast<std::chrono::milliseconds>(r1 - r0).count();
            std::cout << "Reference CPU elapsed time: " << ms_ref << " ms, "
                      << (double)total / (std::max<long long>(ms_ref, 1) * 1000.0) << " MB/s"
                      << std::endl;
            ref_done = true;
          }
          printf("\n############ Validate result ############ \n");
          pass = validate(res, ref);
        }
      }
    }
    if (rc == NGAP_OK && o.write_reports && !o.report_off) {
      // sorted unique reports, one per line as the first 40 are printed
      uint64_t n = 0;
      ngap_get_reports(eng, nullptr, 0, &n);
      std::vector<uint64_t> res(n);
      ngap_get_reports(eng, res.data(), n, &n);
      FILE *f = fopen(o.report_filename.c_str(), "w");
      if (!f) {
        std::cout << BOLDRED << "cannot write " << o.report_filename << RESET << "\n";
        pass = false;
      } else {
        for (uint64_t r : res) fprintf(f, "0x%lx\n", (unsigned long)r);
        fclose(f);
        std::cout << "Reports written: " << res.size() << " to " << o.report_filename << "\n";
      }
    }
    if (pass)
      std::cout << "ngap elapsed time: " << ms / 1000.0 << " seconds, throughput = " << throughput
                << " MB/s " << std::endl;
    else
      std::cout << "ngap elapsed time: " << ms / 1000.0 << " seconds, WRONGTHRPUT(" << throughput
                << ") MB/s " << std::endl;
    return pass ? throughput : -1.0;
  };

  if (o.tuning) {
    // --tuning (ngap.cu:1995-2056): the plain schedule, then the batched
    // engine with the knobs the autotuner picks, keeping the best validated
    // throughput.
    double best = -1;
    std::string best_s;
    auto tt0 = std::chrono::high_resolution_clock::now();
    for (int choice = 0; choice < 2; choice++) {
      ngap_run_opts t = ro;
      std::string desc;
      if (choice == 0) {
        t.schedule = NGAP_SCHED_PLAIN;
        desc = "schedule=plain";
      } else {
        double ms = 0;
        if (ngap_autotune(eng, &ro, 23, &t, &ms) != NGAP_OK) {
          printf("[Tuning] autotuner failed: %s\n", ngap_last_error());
          continue;
        }
        desc = "schedule=batched, window=" + std::to_string(t.window) +
               ", persistent_blocks=" + std::to_string(t.persistent_blocks) +
               ", spill_capacity=" + std::to_string(t.spill_capacity) +
               ", ring_capacity=" + std::to_string(t.ring_capacity) +
               ", root_pass=" + std::to_string(t.root_pass) +
               ", active_threshold=" + std::to_string(t.active_threshold);
      }
      printf("[Tuning] Try choice: \n   %s\n", desc.c_str());
      bool pass = false;
      double thr = one_run(t, pass);
      if (pass && thr > best) {
        best = thr;
        best_s = desc;
        printf("[Tuning] Update best choice: \n   %s max_throughput=%f\n", best_s.c_str(), best);
      } else {
        printf("[Tuning] Keep best choice: \n   %s max_throughput=%f\n", best_s.c_str(), best);
      }
    }
    auto tt1 = std::chrono::high_resolution_clock::now();
    std::cout << "[Tuning] Total time: "
              << std::chrono::duration_cast<std::chrono::microseconds>(tt1 - tt0).count() / 1e6
              << " seconds\n";
    printf("[Tuning] Tuning completed: the best choice is\n   %s\n", best_s.c_str());
  } else {
    bool pass = false;
    one_run(ro, pass);
    if (!pass && o.try_adaptive_aas) {
      // retry with the default worklist spill area and the engine's overflow
      // fallback (work items that overflow again re-run with 16x larger
      // worklists); the reference retries with adaptive AAS batches
      // (ngap.cu:2057-2066)
      printf("Try adaptive aas\n");
      ngap_run_opts t = ro;
      t.spill_capacity = 0;
      t.fallback = 1;
      one_run(t, pass);
    }
  }
  ngap_destroy(eng);
  printf("FINISHED!\n");
  return 0;
}
