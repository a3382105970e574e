// dist_embed: command-line driver of the force-model embedding, shaped like
// the reference's main (cpp/dist_embed.cpp:37-380) and built on the C++
// mirror classes in include/distembed_amd/embedding_algo.hpp.
//
// Reference flags (cpp/dist_embed.cpp:78-133), same spelling and defaults:
//   -input FILE.mtx  -output DIR  -batch B (16384)  -iter T (30)  -lr (0.02)
//   -nsamples ns (5)  -alpha (0)  -beta (0.25)  -dataset NAME  -col_major 0|1
//   -sync_comm 0|1  -fix_batch_training 0|1
// -spmm 1 runs the SpMM sibling as BaselineSpMM::execute does
// (cpp/algo/baseline_spmm.hpp:51-88, batch = all rows, cpp/dist_embed.cpp:179-181):
// each iteration builds a fresh U(-1,1) dense matrix (seed init_seed + i) and
// computes lr * A' X into an output matrix with SpMMAlgo::algo_spmm(1, ...),
// recording one perf_stats entry per iteration; no embedding file is written.
// The algorithms behind -spgemm/-sparse_embedding/-msbfs are not part of this
// library; asking for one exits with status 2.
// Additional flags:
//   -rmat SCALE [-edge_factor 16] [-seed S]     R-MAT graph (a=0.57, b=c=0.19)
//   -sbm N [-blocks 10] [-p_in P] [-p_out P] [-seed S]   stochastic block model
//   -dim D (128; one of 2, 3, 8, 16, 32, 64, 100, 128, 256, 512: each is a
//           compile-time instantiation, as the reference's dimension is; the
//           library itself takes any d in [1, 512])
//   -gather fp32|bf16 (fp32)  -device I (0)  -quirks Q (3 = reference behaviour)
//   -init_seed S (0)
// Outputs, as the reference: DIR/embedding.txt (parrallel_IO.hpp:90-146
// format; DIR is created if missing) and one JSON record appended to
// ./perf_output followed by ",".
#include <distembed_amd/embedding_algo.hpp>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <sys/stat.h>

using namespace distblas::algo;
using namespace distblas::core;
using namespace distblas::io;
using namespace distblas::net;

namespace {

struct Options {
  std::string input_file;
  std::string output_file = "embedding.txt";
  std::string data_set_name;
  int batch_size = 16384;
  double alpha = 0;
  double beta = 0.25;
  int iterations = 30;
  int ns = 5;
  double lr = 0.02;
  bool col_major = false;
  bool sync_comm = false;
  bool fix_batch_training = false;
  bool spmm = false;
  bool other_algo = false;
  // additions
  int rmat_scale = 0;
  int edge_factor = 16;
  int64_t sbm_n = 0;
  int sbm_blocks = 10;
  double p_in = -1, p_out = -1;
  uint64_t seed = 0;
  uint64_t init_seed = 0;
  int dim = 128;
  int gather = FE_GATHER_FP32;
  int device = 0;
  unsigned quirks = FE_QUIRKS_REFERENCE;
};

std::string json_escape(const std::string &s) {
  std::string o;
  for (char c : s) {
    if (c == '"' || c == '\\') o += '\\';
    o += c;
  }
  return o;
}

bool make_dir(const std::string &p) {
  if (p.empty()) return true;
  struct stat st;
  if (stat(p.c_str(), &st) == 0) return S_ISDIR(st.st_mode);
  const size_t slash = p.find_last_of('/');
  if (slash != std::string::npos && slash > 0 && !make_dir(p.substr(0, slash))) return false;
  return mkdir(p.c_str(), 0755) == 0 || errno == EEXIST;
}

void append_record(const Options &o, int batch_size, size_t dim, int world_size,
                   const std::string &stats_json) {
  std::ostringstream js;
  js << "{\n    \"alpha\": " << o.alpha << ",\n    \"algo\": \"Embedding\",\n"
     << "    \"batch_size\": " << batch_size << ",\n    \"beta\": " << o.beta << ",\n"
     << "    \"d\": " << dim << ",\n    \"data_set\": \"" << json_escape(o.data_set_name)
     << "\",\n    \"gather\": \"" << (o.gather == FE_GATHER_BF16 ? "bf16" : "fp32") << "\",\n"
     << "    \"p\": " << world_size << ",\n    \"perf_stats\": " << stats_json
     << ",\n    \"tile_width_fraction\": 1\n}";
  std::ofstream fout("perf_output", std::ios_base::app);
  fout << js.str() << "," << std::endl;
  std::cout << js.str() << std::endl;
}

std::string stats_object(const std::map<std::string, double> &m, const char *indent) {
  std::ostringstream js;
  js << "{";
  bool first = true;
  for (const auto &kv : m) {
    js << (first ? "\n" : ",\n") << indent << "\"" << kv.first << "\": " << kv.second;
    first = false;
  }
  js << "\n" << (indent + 4) << "}";
  return js.str();
}

template <size_t DIM> int run_spmm(const Options &o, SpMat<double> *mat) {
  using INDEX_TYPE = uint64_t;
  using VALUE_TYPE = double;
  auto grid = std::unique_ptr<Process3DGrid>(new Process3DGrid(1, 1, 1, 1));
  grid->device = o.device;
  const int localARows = (int)mat->gRows;
  const int batch_size = localARows; // full batch (cpp/dist_embed.cpp:179-181)
  mat->batch_size = batch_size;
  mat->proc_row_width = localARows;
  mat->proc_col_width = (int)mat->gCols;
  mat->initialize_CSR_blocks(true);
  std::ostringstream arr;
  arr << "[";
  for (int i = 0; i < o.iterations; ++i) {
    auto dense_mat = std::make_shared<DenseMat<INDEX_TYPE, VALUE_TYPE, DIM>>(grid.get(), localARows);
    dense_mat->init_seed = o.init_seed + (uint64_t)i;
    dense_mat->reinitialize();
    auto dense_out = std::make_shared<DenseMat<INDEX_TYPE, VALUE_TYPE, DIM>>(grid.get(), localARows);
    SpMMAlgo<INDEX_TYPE, VALUE_TYPE, DIM> algo(mat, mat, mat, dense_mat.get(), dense_out.get(),
                                               grid.get(), o.alpha, o.beta, o.col_major,
                                               o.sync_comm);
    algo.set_gather_dtype(o.gather);
    algo.set_quirks(o.quirks);
    std::cout << " rank 0 spmm algo started  " << std::endl;
    algo.algo_spmm(1, batch_size, (VALUE_TYPE)o.lr);
    arr << (i ? ",\n        " : "\n        ") << stats_object(algo.perf_stats(), "            ");
  }
  arr << "\n    ]";
  std::cout << " rank 0 SpMM algo completed  " << std::endl;
  append_record(o, batch_size, DIM, 1, arr.str());
  return 0;
}

template <size_t DIM> int run(const Options &o, SpMat<double> *mat) {
  if (o.spmm) return run_spmm<DIM>(o, mat);
  using INDEX_TYPE = uint64_t;
  using VALUE_TYPE = double;
  auto grid = std::unique_ptr<Process3DGrid>(new Process3DGrid(1, 1, 1, 1));
  grid->device = o.device;
  const int world_size = grid->world_size;
  int batch_size = o.batch_size;
  if (o.fix_batch_training) batch_size = batch_size / world_size;
  const int localARows = (int)((mat->gRows + world_size - 1) / world_size);
  const int localBRows = (int)((mat->gCols + world_size - 1) / world_size);
  mat->batch_size = batch_size;
  mat->proc_row_width = localARows;
  mat->proc_col_width = localBRows;
  mat->initialize_CSR_blocks(true);

  auto dense_mat = std::make_shared<DenseMat<INDEX_TYPE, VALUE_TYPE, DIM>>(grid.get(), localARows);
  if (o.init_seed != 0) {
    dense_mat->init_seed = o.init_seed;
    dense_mat->reinitialize();
  }
  // the receiver matrix holds the transposed CSR (cpp/dist_embed.cpp:197-200);
  // with -col_major 1 the step reads it (cpp/algo/algo.hpp:131-133)
  SpMat<double> receiver(grid.get());
  SpMat<double> *recv = mat;
  if (o.col_major) {
    mat->transpose_into(receiver);
    recv = &receiver;
  }
  auto embedding_algo = std::unique_ptr<EmbeddingAlgo<INDEX_TYPE, VALUE_TYPE, DIM>>(
      new EmbeddingAlgo<INDEX_TYPE, VALUE_TYPE, DIM>(mat, recv, mat, dense_mat.get(), grid.get(),
                                                     o.alpha, o.beta, 5, -5, o.col_major,
                                                     o.sync_comm));
  embedding_algo->set_gather_dtype(o.gather);
  embedding_algo->set_quirks(o.quirks);
  std::cout << " rank 0 embedding algo started  " << std::endl;
  embedding_algo->algo_force2_vec_ns(o.iterations, batch_size, o.ns, (VALUE_TYPE)o.lr);
  std::cout << " rank 0 embedding algo completed  " << std::endl;

  if (!make_dir(o.output_file)) {
    std::cerr << "dist_embed: cannot create output directory " << o.output_file << std::endl;
    return 1;
  }
  ParallelIO writer;
  writer.parallel_write(o.output_file + "/embedding.txt", dense_mat->nCoordinates, localARows,
                        (int)DIM, grid.get(), mat);

  append_record(o, batch_size, DIM, world_size,
                stats_object(embedding_algo->perf_stats(), "        "));
  return 0;
}

} // namespace

int main(int argc, char **argv) {
  Options o;
  for (int p = 0; p + 1 < argc; p++) {
    const char *a = argv[p], *v = argv[p + 1];
    if (!strcmp(a, "-input")) o.input_file = v;
    else if (!strcmp(a, "-output")) o.output_file = v;
    else if (!strcmp(a, "-batch")) o.batch_size = atoi(v);
    else if (!strcmp(a, "-iter")) o.iterations = atoi(v);
    else if (!strcmp(a, "-alpha")) o.alpha = atof(v);
    else if (!strcmp(a, "-lr")) o.lr = atof(v);
    else if (!strcmp(a, "-nsamples")) o.ns = atoi(v);
    else if (!strcmp(a, "-beta")) o.beta = atof(v);
    else if (!strcmp(a, "-dataset")) o.data_set_name = v;
    else if (!strcmp(a, "-col_major")) o.col_major = atoi(v) != 0;
    else if (!strcmp(a, "-sync_comm")) o.sync_comm = atoi(v) != 0;
    else if (!strcmp(a, "-fix_batch_training")) o.fix_batch_training = atoi(v) == 1;
    else if (!strcmp(a, "-spmm")) o.spmm = atoi(v) == 1;
    else if (!strcmp(a, "-spgemm") || !strcmp(a, "-sparse_embedding") || !strcmp(a, "-msbfs") ||
             !strcmp(a, "-save_results"))
      o.other_algo = o.other_algo || atoi(v) == 1;
    else if (!strcmp(a, "-rmat")) o.rmat_scale = atoi(v);
    else if (!strcmp(a, "-edge_factor")) o.edge_factor = atoi(v);
    else if (!strcmp(a, "-sbm")) o.sbm_n = atoll(v);
    else if (!strcmp(a, "-blocks")) o.sbm_blocks = atoi(v);
    else if (!strcmp(a, "-p_in")) o.p_in = atof(v);
    else if (!strcmp(a, "-p_out")) o.p_out = atof(v);
    else if (!strcmp(a, "-seed")) o.seed = strtoull(v, nullptr, 10);
    else if (!strcmp(a, "-init_seed")) o.init_seed = strtoull(v, nullptr, 10);
    else if (!strcmp(a, "-dim")) o.dim = atoi(v);
    else if (!strcmp(a, "-gather")) o.gather = !strcmp(v, "bf16") ? FE_GATHER_BF16 : FE_GATHER_FP32;
    else if (!strcmp(a, "-device")) o.device = atoi(v);
    else if (!strcmp(a, "-quirks")) o.quirks = (unsigned)atoi(v);
  }
  if (o.other_algo) {
    std::cerr << "dist_embed: only the force-model embedding and -spmm are built (no -spgemm, "
                 "-sparse_embedding, -msbfs or -save_results)" << std::endl;
    return 2;
  }
  Process3DGrid io_grid(1, 1, 1, 1);
  SpMat<double> mat(&io_grid);
  try {
    if (o.rmat_scale > 0 || o.sbm_n > 0) {
      fe_graph g{};
      fe_status st;
      if (o.rmat_scale > 0) {
        st = fe_gen_rmat(o.rmat_scale, o.edge_factor, 0.57, 0.19, 0.19, o.seed, &g);
      } else {
        const double nb = (double)o.sbm_n / o.sbm_blocks;
        const double p_in = o.p_in >= 0 ? o.p_in : std::min(1.0, 18.0 / (nb - 1));
        const double p_out = o.p_out >= 0 ? o.p_out : std::min(1.0, 2.0 / ((double)o.sbm_n - nb));
        st = fe_gen_sbm(o.sbm_n, o.sbm_blocks, p_in, p_out, o.seed, &g);
      }
      if (st != FE_OK) {
        std::cerr << "dist_embed: graph generation failed: " << fe_last_error(nullptr) << std::endl;
        return 1;
      }
      mat.set_csr(g.n, g.row_ptr, g.col_idx);
      fe_free_graph(&g);
    } else {
      if (o.input_file.empty()) {
        std::cerr << "dist_embed: give -input FILE.mtx, -rmat SCALE or -sbm N" << std::endl;
        return 2;
      }
      std::cout << " rank 0 reading data from file path:  " << o.input_file << std::endl;
      ParallelIO reader;
      reader.parallel_read_MM<int64_t, int, double>(o.input_file, &mat, true);
    }
    std::cout << " rank 0 gROWs  " << mat.gRows << "gCols" << mat.gCols << std::endl;
    switch (o.dim) {
    case 2: return run<2>(o, &mat);
    case 3: return run<3>(o, &mat);
    case 8: return run<8>(o, &mat);
    case 16: return run<16>(o, &mat);
    case 32: return run<32>(o, &mat);
    case 64: return run<64>(o, &mat);
    case 100: return run<100>(o, &mat);
    case 128: return run<128>(o, &mat);
    case 256: return run<256>(o, &mat);
    case 512: return run<512>(o, &mat);
    default:
      std::cerr << "dist_embed: -dim must be one of 2, 3, 8, 16, 32, 64, 100, 128, 256, 512"
                << std::endl;
      return 2;
    }
  } catch (const std::exception &e) {
    std::cerr << "dist_embed: " << e.what() << std::endl;
    return 1;
  }
}
