// Checkpoint save/load (TeNeS src/iTPS/saveload_tensors.cpp:38-230).
//
// Directory layout and params.dat as TeNeS writes them.  The tensor files
// themselves are TeNeS's mptensor Tensor::save output there; that library is
// not part of this build, so a tensor file here is
//   8 bytes "TNSRF64\0", int64 rank, int64 shape[rank], float64 data
// (row-major, little-endian), and the bond-weight files carry full precision
// (17 significant digits) where TeNeS's stream default writes 6.
#include "checkpoint.hpp"

#include <sys/stat.h>

#include <cerrno>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>

namespace tenes {
namespace itps {

namespace {

const char *const kStem[kCheckpointKinds] = {"C1", "C2", "C3", "C4", "Et",
                                             "Er", "Eb", "El", "T"};
constexpr int kTn = 8;
constexpr char kMagic[8] = {'T', 'N', 'S', 'R', 'F', '6', '4', '\0'};

std::string file_of(const std::string &dir, int kind, int site) {
  return dir + "/" + kStem[kind] + "_" + std::to_string(site) + ".dat";
}

bool is_dir(const std::string &path) {
  struct stat sb;
  return stat(path.c_str(), &sb) == 0 && S_ISDIR(sb.st_mode);
}

// The text before a '#' of a params.dat line (util::drop_comment).
std::string drop_comment(const std::string &line) {
  const auto p = line.find('#');
  return p == std::string::npos ? line : line.substr(0, p);
}

void write_tensor(const std::string &path, const tn::Tensor &t) {
  std::vector<double> h((size_t)t.size());
  if (!h.empty()) tn::to_host(t, h.data());
  std::ofstream f(path, std::ios::binary);
  TN_REQUIRE(f.good(), "cannot write " + path);
  const int64_t r = t.rank();
  f.write(kMagic, sizeof(kMagic));
  f.write(reinterpret_cast<const char *>(&r), sizeof(r));
  f.write(reinterpret_cast<const char *>(t.shape.data()), sizeof(int64_t) * r);
  f.write(reinterpret_cast<const char *>(h.data()), sizeof(double) * h.size());
  TN_REQUIRE(f.good(), "write failed: " + path);
}

tn::Tensor read_tensor(tn::Context &ctx, const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  TN_REQUIRE(f.good(), "cannot read " + path);
  char magic[sizeof(kMagic)];
  int64_t r = -1;
  f.read(magic, sizeof(magic));
  f.read(reinterpret_cast<char *>(&r), sizeof(r));
  TN_REQUIRE(f.good() && std::memcmp(magic, kMagic, sizeof(kMagic)) == 0,
             path + " is not a tensor file of this library");
  TN_REQUIRE(r >= 0 && r <= 16, path + ": bad rank");
  tn::Shape shape((size_t)r);
  f.read(reinterpret_cast<char *>(shape.data()), sizeof(int64_t) * r);
  int64_t n = 1;
  for (int64_t d : shape) {
    TN_REQUIRE(f.good() && d >= 0 && d < (int64_t(1) << 31), path + ": bad shape");
    n *= d;
  }
  std::vector<double> h((size_t)n);
  f.read(reinterpret_cast<char *>(h.data()), sizeof(double) * n);
  TN_REQUIRE(f.good(), path + ": truncated data");
  return tn::from_host(ctx, shape, h.data());
}

// resize_tensor (src/tensor.cpp:114-148): zero padding first, then slicing.
tn::Tensor resize_to(const tn::Tensor &src, const tn::Shape &target) {
  tn::Shape ext = src.shape;
  bool extend = false;
  for (size_t i = 0; i < ext.size(); ++i)
    if (ext[i] < target[i]) {
      ext[i] = target[i];
      extend = true;
    }
  tn::Tensor a = extend ? tn::extend(src, ext) : src;
  for (size_t i = 0; i < ext.size(); ++i)
    if (a.shape[i] > target[i]) a = tn::slice(a, (int)i, 0, target[i]);
  return a;
}

}  // namespace

void save_tensors(const CheckpointState &st, const std::string &dir, int chi) {
  TN_REQUIRE(!dir.empty(), "save_tensors: empty directory name");
  if (!is_dir(dir))
    TN_REQUIRE(mkdir(dir.c_str(), 0755) == 0 || errno == EEXIST,
               "save_tensors: cannot create " + dir);
  const int N = st.n_unit;
  for (int k = 0; k < kCheckpointKinds; ++k)
    for (int i = 0; i < N; ++i)
      TN_REQUIRE(st.arr[k][i].defined(),
                 std::string("save_tensors: ") + kStem[k] + "[" + std::to_string(i) +
                     "] is not set (initialise the environment first)");
  if (chi <= 0) chi = (int)st.arr[0][0].shape[0];
  {
    std::ofstream ofs(dir + "/params.dat");
    TN_REQUIRE(ofs.good(), "cannot write " + dir + "/params.dat");
    ofs << 1 << " # Format_Version\n";
    ofs << N << " # N_UNIT\n";
    ofs << chi << " # CHI\n";
    for (int i = 0; i < N; ++i) {
      const tn::Tensor &T = st.arr[kTn][i];
      TN_REQUIRE(T.rank() >= 5, "save_tensors: site tensors have at least 5 legs");
      for (int j = 0; j < 4; ++j) ofs << T.shape[j] << " ";
      ofs << T.shape[4] << " # Shape of Tn[" << i << "]\n";
    }
    TN_REQUIRE(ofs.good(), "write failed: " + dir + "/params.dat");
  }
  for (int i = 0; i < N; ++i)
    for (int k = kCheckpointKinds - 1; k >= 0; --k)
      write_tensor(file_of(dir, k, i), st.arr[k][i]);
  for (int i = 0; i < N; ++i) {
    std::ofstream ofs(dir + "/lambda_" + std::to_string(i) + ".dat");
    TN_REQUIRE(ofs.good(), "cannot write bond weights in " + dir);
    ofs << std::setprecision(17);
    for (int j = 0; j < 4; ++j)
      for (int64_t k = 0; k < st.arr[kTn][i].shape[j]; ++k) {
        const auto &l = (*st.lambda)[i][j];
        ofs << ((size_t)k < l.size() ? l[k] : 1.0) << "\n";
      }
  }
}

void load_tensors(CheckpointState &st, tn::Context &ctx, const std::string &dir,
                  int chi) {
  if (!is_dir(dir)) throw tn::Error(dir + " does not exist.");
  const int N = st.n_unit;
  std::ifstream ifs(dir + "/params.dat");
  if (!ifs.good())
    throw tn::Error("load_tensors: " + dir +
                    "/params.dat is missing (saved tensor format version 0 is not supported)");
  std::string line;
  std::getline(ifs, line);
  const int version = std::stoi(drop_comment(line));
  if (version != 1)
    throw tn::Error("ERROR: Unknown saved tensor format version: " + std::to_string(version));
  std::getline(ifs, line);
  const int loaded_n = std::stoi(drop_comment(line));
  if (loaded_n != N)
    throw tn::Error("ERROR: N_UNIT is " + std::to_string(N) + " but loaded N_UNIT has " +
                    std::to_string(loaded_n));
  std::getline(ifs, line);  // CHI of the saved run (a differing CHI is allowed)
  std::vector<std::vector<int64_t>> loaded_shape(N, std::vector<int64_t>(5));
  for (int i = 0; i < N; ++i) {
    std::getline(ifs, line);
    std::istringstream ss(drop_comment(line));
    for (int j = 0; j < 5; ++j) ss >> loaded_shape[i][j];
    TN_REQUIRE(!ss.fail(), "load_tensors: bad shape line in params.dat");
    const tn::Tensor &T = st.arr[kTn][i];
    if (T.defined() && T.rank() >= 5 && T.shape[4] != loaded_shape[i][4])
      throw tn::Error("ERROR: dimension of the physical bond of the tensor " +
                      std::to_string(i) + " is " + std::to_string(T.shape[4]) +
                      " but loaded tensor has " + std::to_string(loaded_shape[i][4]));
  }
  // read everything first: a failure leaves the environment untouched
  std::vector<tn::Tensor> next[kCheckpointKinds];
  for (int k = 0; k < kCheckpointKinds; ++k) {
    next[k].resize(N);
    for (int i = 0; i < N; ++i) {
      tn::Tensor t = read_tensor(ctx, file_of(dir, k, i));
      const tn::Tensor &cur = st.arr[k][i];
      tn::Shape target = t.shape;
      if (cur.defined()) {
        if (cur.rank() != t.rank())
          throw tn::Error(std::string("ERROR: rank mismatch in load_tensor: ") + kStem[k] +
                          "[" + std::to_string(i) + "] has " + std::to_string(cur.rank()) +
                          " legs, but loaded one has " + std::to_string(t.rank()) +
                          " legs.  HINT: check the calculation mode: the number of legs "
                          "differs between ground state and finite temperature.");
        target = cur.shape;
      } else if (chi > 0 && k < kTn) {
        target[0] = target[1] = chi;  // the two environment legs of C and eT
      }
      next[k][i] = resize_to(t, target);
    }
  }
  std::vector<std::vector<std::vector<double>>> lam(N, std::vector<std::vector<double>>(4));
  for (int i = 0; i < N; ++i) {
    std::ifstream lf(dir + "/lambda_" + std::to_string(i) + ".dat");
    TN_REQUIRE(lf.good(), "load_tensors: missing lambda_" + std::to_string(i) + ".dat");
    for (int j = 0; j < 4; ++j) {
      for (int64_t k = 0; k < loaded_shape[i][j]; ++k) {
        double v = 0.0;
        lf >> v;
        TN_REQUIRE(!lf.fail(), "load_tensors: truncated lambda_" + std::to_string(i) + ".dat");
        lam[i][j].push_back(v);
      }
      lam[i][j].resize((size_t)next[kTn][i].shape[j], 0.0);
    }
  }
  for (int k = 0; k < kCheckpointKinds; ++k) st.arr[k] = next[k];
  *st.lambda = lam;
}

}  // namespace itps
}  // namespace tenes
