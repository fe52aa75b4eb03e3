// entropical_main.cpp -- command-line front end of the MI355X KDE path.
//
// Same modes, options, messages and output layout as the reference's
// entropical program (entropical.cpp:25-240 and the mode drivers), on the
// batched device drivers of libentropical_kde: every mode computes all of its
// densities in one engine batch on the GPU and reduces them there.
//
//   modes:   transs autoent kldiv mi dens negs amise thumb
//   options: -i/--input FILE   coordinates, column-oriented whitespace-separated text,
//                              or a GROMACS .xtc trajectory (row = frame, column 3a+e)
//            -o/--output FILE  output file (default: stdout)
//            -c/--columns "i j ..."  1-based column selection (default: all)
//            -H/--bandwidths FILE   bandwidth per selected column, whitespace-separated
//            -t/--tau N        lag (transs, autoent; default 1)
//            -B/--bits         log2 instead of ln
//            -D/--dims N       kernel dimension for dens (1, 2 or 3; default 1)
//            --taus "a b [c]"  lags of a dens kernel (dims 2 or 3)
//            -v/--verbose, -h/--help
// Environment: ENT_DEVICE selects the GPU (default 0), ENT_METHOD=brute the
// all-pairs search (default: cell lists).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "entropical_kde.h"

namespace {

[[noreturn]] void die(const std::string& msg) {
  std::cerr << msg << std::endl;
  std::exit(EXIT_FAILURE);
}

const char* kGeneralHelp =
    "\n"
    " entropical (MI355X)\n\n"
    " compute entropy-based/information measures like\n"
    " local probabilities, mutual information, transfer entropies, etc.\n"
    " code is accelerated by HIP kernels on an AMD Instinct MI355X GPU.\n\n"
    "modes:\n"
    "  transs:    compute transfer entropies\n"
    "  autoent:   compute conditional entropy on own past\n"
    "  kldiv:     compute Kullback-Leibler divergences\n"
    "  mi:        compute mutual informations\n"
    "  dens:      compute local probability densities in 1, 2 or 3 dimensions\n"
    "  negs:      compute negentropies\n"
    "  amise:     estimate bandwidths from AMISE-minimization\n"
    "  thumb:     estimate bandwidths from rule-of-thumb\n\n"
    "usage:\n"
    "  entropical MODE --option1 --option2 ...\n\n"
    "for a list of available options per mode, run with '-h' option, e.g.\n"
    "  entropical transs -h\n\n";

// ------------------------------------------------------------- options
struct Opt {
  std::string long_name;
  char short_name;
  bool is_switch;
  std::string def;
  std::string help;
};

struct Mode {
  std::string title;
  std::vector<Opt> opts;
};

std::vector<Opt> common_opts() {
  return {
      {"help", 'h', true, "", "show this help."},
      {"input", 'i', false, "", "coordinates (column-oriented space-separated plain text; required)."},
      {"output", 'o', false, "", "output file (default: stdout)"},
      {"columns", 'c', false, "",
       "column selection, space separated indices;\ne.g. -c \"1 3 5\" for first, third and fifth columns\n"
       "(default: read all)"},
      {"verbose", 'v', true, "", "verbose output."},
  };
}

std::map<std::string, Mode> modes() {
  const Opt bw{"bandwidths", 'H', false, "", "bandwidths for univariate density estimation as space separated values."};
  const Opt bits{"bits", 'B', true, "", "use log2 instead of ln"};
  const Opt tau{"tau", 't', false, "1", "lagtime in # frames, default: 1"};
  std::map<std::string, Mode> m;
  auto mk = [&](const std::string& title, std::vector<Opt> extra) {
    Mode md;
    md.title = title;
    md.opts = common_opts();
    md.opts.insert(md.opts.end(), extra.begin(), extra.end());
    return md;
  };
  m["transs"] = mk("transs - compute transfer entropies between time series", {tau, bw, bits});
  m["autoent"] = mk("autoent - compute conditional entropy against own past", {tau, bw, bits});
  m["kldiv"] = mk("kldiv - compute Kullback-Leibler divergences between time series", {bw, bits});
  m["mi"] = mk("mi - compute mutual informations between time series", {bw, bits});
  m["dens"] = mk("dens - compute local probability densities of time series",
                 {bw,
                  {"taus", 0, false, "",
                   "lag times for probability density estimation (must be equal to # of dimensions; default: no "
                   "time lag.)"},
                  {"dims", 'D', false, "1", "kernel dimensionality for multivariate density estimation (dens)"}});
  m["negs"] = mk("negs - compute negentropies of time series", {bw, bits});
  m["amise"] = mk("amise - estimate probability kernel bandwidths via amise-minimization", {});
  m["thumb"] = mk("thumb - estimate probability kernel bandwidths via rule-of-thumb", {});
  return m;
}

void print_help(const Mode& md, std::ostream& os) {
  os << md.title << ":\n";
  for (const Opt& o : md.opts) {
    std::string lhs = "  ";
    if (o.short_name) lhs += std::string("-") + o.short_name + " [ --" + o.long_name + " ]";
    else lhs += "--" + o.long_name;
    if (!o.is_switch) lhs += " arg";
    if (!o.def.empty()) lhs += " (=" + o.def + ")";
    os << lhs;
    std::istringstream hs(o.help);
    std::string line;
    bool first = true;
    while (std::getline(hs, line)) {
      if (first) os << std::string(lhs.size() < 34 ? 34 - lhs.size() : 1, ' ') << line << "\n";
      else os << std::string(34, ' ') << line << "\n";
      first = false;
    }
  }
}

// Parses argv[2..] against the mode's options: --name value, --name=value,
// -x value, -xvalue, and switches.  Unknown options are an error.
std::map<std::string, std::string> parse(const Mode& md, int argc, char** argv) {
  std::map<std::string, std::string> v;
  for (const Opt& o : md.opts) v[o.long_name] = o.is_switch ? "0" : o.def;
  std::set<std::string> given;
  for (int a = 2; a < argc; ++a) {
    std::string s = argv[a];
    const Opt* opt = nullptr;
    std::string value;
    bool has_value = false;
    if (s.rfind("--", 0) == 0) {
      std::string name = s.substr(2);
      size_t eq = name.find('=');
      if (eq != std::string::npos) {
        value = name.substr(eq + 1);
        name = name.substr(0, eq);
        has_value = true;
      }
      for (const Opt& o : md.opts)
        if (o.long_name == name) opt = &o;
    } else if (s.size() >= 2 && s[0] == '-') {
      for (const Opt& o : md.opts)
        if (o.short_name && o.short_name == s[1]) opt = &o;
      if (s.size() > 2) {
        value = s.substr(2);
        has_value = true;
      }
    }
    if (!opt) throw std::runtime_error("unrecognised option '" + s + "'");
    if (given.count(opt->long_name))
      throw std::runtime_error("option '" + opt->long_name + "' cannot be specified more than once");
    given.insert(opt->long_name);
    if (opt->is_switch) {
      if (has_value) throw std::runtime_error("option '" + opt->long_name + "' does not take any arguments");
      v[opt->long_name] = "1";
      continue;
    }
    if (!has_value) {
      if (a + 1 >= argc) throw std::runtime_error("the required argument for option '--" + opt->long_name + "' is missing");
      value = argv[++a];
    }
    v[opt->long_name] = value;
  }
  return v;
}

unsigned long to_ulong(const std::string& s, const std::string& what) {
  errno = 0;
  char* end = nullptr;
  unsigned long x = std::strtoul(s.c_str(), &end, 10);
  if (s.empty() || *end != '\0' || errno) throw std::runtime_error("the argument ('" + s + "') for option '" + what + "' is invalid");
  return x;
}

std::vector<std::string> split_ws(const std::string& s) {
  std::istringstream is(s);
  std::vector<std::string> out;
  std::string t;
  while (is >> t) out.push_back(t);
  return out;
}

// ------------------------------------------------------------------- I/O
// Tools::IO::read_coords (tools.hxx:154-233) with primary index 'C': the
// number of columns comes from the first non-empty line, the number of rows
// is the number of non-empty lines, selected columns are read in file order.
// A line with a different number of values is an error here (the reference
// reads on across line ends).
struct Coords {
  std::vector<float> x;  // column-major, n_cols * n_rows
  size_t n_rows = 0, n_cols = 0;
  std::vector<size_t> selected;  // 1-based, as given
};

bool is_xtc(const std::string& fname) {
  return fname.size() > 4 && fname.compare(fname.size() - 4, 4, ".xtc") == 0;
}

Coords read_coords(const std::string& fname, const std::string& columns) {
  Coords C;
  for (const std::string& c : split_ws(columns)) {
    errno = 0;
    char* end = nullptr;
    long long v = std::strtoll(c.c_str(), &end, 10);
    if (*end != '\0' || errno) die("error: specify column in range of [1, N_COLUMNS]!");
    if (v <= 0) die("error: specify column in range of [1, N_COLUMNS]!");
    C.selected.push_back((size_t)v);
  }
  // the whole file, column-major: full[c * n_rows + r]
  std::vector<float> full;
  size_t n_rows = 0, n_file_cols = 0;
  if (is_xtc(fname)) {
    // GROMACS trajectory: frame = row, coordinate e of atom a = column 3a+e
    // (CoordsFile::XtcHandler, coords_file/coords_file.cpp:95-114)
    size_t nf = 0;
    uint32_t na = 0;
    if (ent_read_xtc(fname.c_str(), nullptr, 0, &nf, &na) != ENT_OK) die(std::string("error: ") + ent_last_error());
    full.resize(nf * 3 * (size_t)na);
    if (ent_read_xtc(fname.c_str(), full.data(), full.size(), &nf, &na) != ENT_OK)
      die(std::string("error: ") + ent_last_error());
    n_rows = nf;
    n_file_cols = 3 * (size_t)na;
  } else {
    std::ifstream ifs(fname);
    if (!ifs) die("error: cannot open input file '" + fname + "'");
    std::vector<std::vector<float>> rows;
    std::string line;
    while (std::getline(ifs, line)) {
      if (line.empty()) continue;
      std::vector<float> r;
      const char* p = line.c_str();
      char* end = nullptr;
      for (;;) {
        while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
        if (!*p) break;
        float f = std::strtof(p, &end);
        if (end == p) die("error: non-numeric value in input file '" + fname + "'");
        r.push_back(f);
        p = end;
      }
      if (r.empty()) continue;
      if (rows.empty()) n_file_cols = r.size();
      if (r.size() != n_file_cols)
        die("error: row " + std::to_string(rows.size() + 1) + " of '" + fname + "' has " + std::to_string(r.size()) +
            " values, expected " + std::to_string(n_file_cols));
      rows.push_back(std::move(r));
    }
    n_rows = rows.size();
    full.resize(n_rows * n_file_cols);
    for (size_t r = 0; r < n_rows; ++r)
      for (size_t c = 0; c < n_file_cols; ++c) full[c * n_rows + r] = rows[r][c];
  }
  std::vector<size_t> use;  // 0-based file columns, file order
  if (C.selected.empty()) {
    for (size_t i = 0; i < n_file_cols; ++i) {
      use.push_back(i);
      C.selected.push_back(i + 1);
    }
  } else {
    std::set<size_t> s;
    for (size_t c : C.selected) {
      if (n_file_cols <= c - 1)
        die("error: given column index is not compatible with available number of columns (" +
            std::to_string(n_file_cols) + ")");
      if (!s.insert(c - 1).second) die("error: column " + std::to_string(c) + " selected twice");
    }
    use.assign(s.begin(), s.end());
  }
  C.n_rows = n_rows;
  C.n_cols = use.size();
  C.x.resize(C.n_rows * C.n_cols);
  for (size_t k = 0; k < C.n_cols; ++k)
    std::memcpy(C.x.data() + k * n_rows, full.data() + use[k] * n_rows, sizeof(float) * n_rows);
  return C;
}

// selected_coords_bandwidths (tools.hxx:275-305): whitespace-separated values
// from a file, one per selected column.
std::vector<float> read_bandwidths(const std::string& fname, size_t n_cols) {
  std::vector<float> hs;
  std::ifstream ifs(fname);
  float b;
  while (ifs >> b) hs.push_back(b);
  if (hs.size() != n_cols)
    die("error: number of given bandwidth values does not match the number of selected columns!");
  return hs;
}

// ------------------------------------------------------------------ GPU
struct Gpu {
  ent_ctx* ctx = nullptr;
  float* X = nullptr;
  int method = ENT_METHOD_CELL;
  ~Gpu() {
    if (ctx && X) ent_dev_free(ctx, X);
    if (ctx) ent_ctx_destroy(ctx);
  }
};

void check(int rc) {
  if (rc != ENT_OK) die(std::string("error: ") + ent_last_error());
}

void upload(Gpu& g, const Coords& C) {
  int dev = 0;
  if (const char* d = std::getenv("ENT_DEVICE")) dev = std::atoi(d);
  if (const char* m = std::getenv("ENT_METHOD")) g.method = std::strcmp(m, "brute") == 0 ? ENT_METHOD_BRUTE : ENT_METHOD_CELL;
  check(ent_ctx_create(dev, &g.ctx));
  check(ent_dev_alloc(g.ctx, sizeof(float) * C.x.size(), (void**)&g.X));
  check(ent_h2d(g.ctx, g.X, C.x.data(), sizeof(float) * C.x.size()));
}

// Tools::sum1_normalized (tools.hxx:33-43), compensated sum.
void sum1_normalize(std::vector<double>& p) {
  double s = 0.0, c = 0.0;
  for (double v : p) {
    double y = v - c, t = s + y;
    c = (t - s) - y;
    s = t;
  }
  if (s != 0.0)
    for (double& v : p) v /= s;
}

// Densities of a list of jobs over the uploaded columns (one batch).
std::vector<std::vector<double>> batch_densities(Gpu& g, size_t n_rows, const std::vector<std::vector<unsigned>>& cols,
                                                 const std::vector<std::vector<float>>& hs,
                                                 const std::vector<unsigned>& taus) {
  unsigned tmax = 0;
  for (unsigned t : taus) tmax = t > tmax ? t : tmax;
  const size_t n = n_rows - tmax;
  std::vector<ent_job> jobs(cols.size());
  for (size_t j = 0; j < cols.size(); ++j) {
    ent_job& J = jobs[j];
    std::memset(&J, 0, sizeof(J));
    J.n_dim = (uint32_t)cols[j].size();
    J.n = (uint32_t)n;
    J.n_rows = (uint32_t)n_rows;
    for (size_t e = 0; e < cols[j].size(); ++e) {
      J.off[e] = (uint64_t)cols[j][e] * n_rows + (taus.empty() ? 0 : taus[e]);
      J.h[e] = hs[j][e];
    }
    J.out_off = (uint64_t)j * n;
  }
  double* d = nullptr;
  check(ent_dev_alloc(g.ctx, sizeof(double) * n * jobs.size(), (void**)&d));
  check(ent_batch_densities_dev(g.ctx, g.X, jobs.data(), jobs.size(), g.method, d));
  std::vector<double> all(n * jobs.size());
  check(ent_d2h(g.ctx, all.data(), d, sizeof(double) * all.size()));
  check(ent_dev_free(g.ctx, d));
  std::vector<std::vector<double>> out(jobs.size());
  for (size_t j = 0; j < jobs.size(); ++j) out[j].assign(all.begin() + j * n, all.begin() + (j + 1) * n);
  return out;
}

void print_matrix(std::ostream& os, const std::vector<double>& m, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    for (size_t j = 0; j < n; ++j) os << " " << m[i * n + j];
    os << "\n";
  }
}

// ----------------------------------------------------------------- modes
// The reference reads --input without a default; a missing one is an
// argument-parsing error there (entropical.cpp:226-231).
struct ParseError : std::runtime_error {
  ParseError() : std::runtime_error("error parsing arguments!") {}
};

int run(const std::string& mode, std::map<std::string, std::string>& a) {
  if (a["input"].empty()) throw ParseError();
  std::ofstream ofs;
  if (!a["output"].empty()) {
    ofs.open(a["output"]);
    if (!ofs) die("error: cannot open output file '" + a["output"] + "'");
  }
  std::ostream& out = a["output"].empty() ? std::cout : ofs;
  const bool bits = a.count("bits") && a["bits"] == "1";
  Coords C = read_coords(a["input"], a["columns"]);
  const size_t n = C.n_cols, N = C.n_rows;
  if (N == 0 || n == 0) die("error: no coordinates in input file '" + a["input"] + "'");
  Gpu g;
  if (mode == "amise" || mode == "thumb") {
    upload(g, C);
    std::vector<float> h(n);
    check(ent_hestimate_dev(g.ctx, g.X, N, n, mode == "amise" ? ENT_HEST_AMISE : ENT_HEST_THUMB, h.data()));
    for (float v : h) out << " " << v;
    out << std::endl;
    return EXIT_SUCCESS;
  }
  std::vector<float> hs = read_bandwidths(a["bandwidths"], n);
  if (mode == "mi" || mode == "kldiv") {
    upload(g, C);
    std::vector<double> m(n * n);
    if (mode == "mi") {
      check(ent_mi_matrix_dev(g.ctx, g.X, N, n, hs.data(), bits, g.method, m.data()));
      for (double& v : m) v = (float)v;  // mi.cpp:36 keeps the matrix in float
    } else {
      check(ent_kldiv_dev(g.ctx, g.X, N, n, hs.data(), bits, g.method, m.data()));
    }
    print_matrix(out, m, n);
  } else if (mode == "transs") {
    const uint32_t tau = (uint32_t)to_ulong(a["tau"], "--tau");
    upload(g, C);
    std::vector<double> m(n * n);
    check(ent_te_matrix_dev(g.ctx, g.X, N, n, hs.data(), &tau, 1, bits, g.method, m.data()));
    print_matrix(out, m, n);
  } else if (mode == "autoent") {
    const uint32_t tau = (uint32_t)to_ulong(a["tau"], "--tau");
    upload(g, C);
    std::vector<double> v(n);
    check(ent_autoent_dev(g.ctx, g.X, N, n, hs.data(), tau, bits, g.method, v.data()));
    for (double x : v) out << " " << x << "\n";
  } else if (mode == "negs") {
    upload(g, C);
    std::vector<double> v(n);
    check(ent_negs_dev(g.ctx, g.X, N, n, hs.data(), bits, g.method, v.data()));
    for (double x : v) out << " " << (float)x;
    out << "\n";
  } else if (mode == "dens") {
    // Dens::main (dens.cpp:110-187)
    const unsigned long dims = to_ulong(a["dims"], "--dims");
    if (dims == 0 || dims > 3) die("error: only allowed kernel dimensions are 1, 2 or 3.");
    if (dims == 1) {
      if (!a["taus"].empty())
        die("error: definition of lagtimes (taus) for 1d probability density does not make sense.");
      upload(g, C);
      std::vector<std::vector<unsigned>> cols;
      std::vector<std::vector<float>> h;
      for (unsigned j = 0; j < n; ++j) {
        cols.push_back({j});
        h.push_back({hs[j]});
      }
      auto d = batch_densities(g, N, cols, h, {});
      for (auto& v : d) sum1_normalize(v);
      for (size_t i = 0; i < N; ++i) {
        for (size_t j = 0; j < n; ++j) out << " " << d[j][i];
        out << "\n";
      }
    } else {
      std::vector<unsigned> taus(dims, 0);
      std::vector<std::string> ts = split_ws(a["taus"]);
      if (!ts.empty()) {
        if (ts.size() != dims) die("error: number of lagtime parameters does not match specified kernel dimension!");
        for (size_t i = 0; i < dims; ++i) taus[i] = (unsigned)to_ulong(ts[i], "--taus");
      }
      for (unsigned t : taus)
        if (t >= N) die("error: lag time (tau) must be smaller than the number of rows");
      std::vector<std::vector<unsigned>> cols;
      std::vector<std::vector<float>> h;
      std::vector<std::string> labels;
      for (unsigned i = 0; i < n; ++i)
        for (unsigned j = i + 1; j < n; ++j) {
          if (dims == 2) {
            cols.push_back({i, j});
            h.push_back({hs[i], hs[j]});
            labels.push_back(std::to_string(C.selected[i]) + "_" + std::to_string(C.selected[j]));
          } else {
            for (unsigned k = j + 1; k < n; ++k) {
              cols.push_back({i, j, k});
              h.push_back({hs[i], hs[j], hs[k]});
              labels.push_back(std::to_string(C.selected[i]) + "_" + std::to_string(C.selected[j]) + "_" +
                               std::to_string(C.selected[k]));
            }
          }
        }
      if (cols.empty()) return EXIT_SUCCESS;
      upload(g, C);
      auto d = batch_densities(g, N, cols, h, taus);
      for (size_t j = 0; j < d.size(); ++j) {
        sum1_normalize(d[j]);
        for (double v : d[j]) out << labels[j] << " " << v << "\n";
      }
    }
  }
  return EXIT_SUCCESS;
}

}  // namespace

int main(int argc, char** argv) {
  std::map<std::string, Mode> M = modes();
  if (argc <= 2) {
    std::cerr << kGeneralHelp;
    return EXIT_FAILURE;
  }
  const std::string mode = argv[1];
  if (!M.count(mode)) {
    std::cerr << "\nerror: unrecognized mode '" << mode << "'\n\n" << kGeneralHelp;
    return EXIT_FAILURE;
  }
  std::map<std::string, std::string> args;
  try {
    args = parse(M[mode], argc, argv);
    if (args["help"] == "1") {
      print_help(M[mode], std::cout);
      std::cout << std::endl;
      return EXIT_SUCCESS;
    }
    return run(mode, args);
  } catch (const ParseError& e) {
    std::cerr << "\n" << e.what() << "\n\n";
    print_help(M[mode], std::cerr);
    std::cerr << std::endl;
    return EXIT_FAILURE;
  } catch (const std::exception& e) {
    std::cerr << "\n" << e.what() << "\n\n";
    print_help(M[mode], std::cerr);
    std::cerr << std::endl;
    return EXIT_FAILURE;
  }
}
