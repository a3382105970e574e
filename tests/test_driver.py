"""The C++ mirror of the reference API (include/distembed_amd/embedding_algo.hpp)
and the dist_embed driver built on it, without a GPU: the reference's driver
code shape compiles against the mirror, and the driver's error paths."""
import subprocess
from pathlib import Path

import pytest

from distembed_amd import _lib, build

ROOT = Path(__file__).resolve().parent.parent
DRIVER = build.DRIVER

# The calls of cpp/dist_embed.cpp:155-347 that drive the embedding, in the
# reference's own qualified names and argument order.
REFERENCE_SHAPED = r"""
#include <distembed_amd/embedding_algo.hpp>
#include <memory>
#include <string>
using namespace distblas::io;
using namespace distblas::core;
using INDEX_TYPE = uint64_t;
using VALUE_TYPE = double;
int main(int argc, char **argv) {
  const int dimension = DIMENSION;
  int world_size = 1, batch_size = 16384, iterations = 30, ns = 5;
  double alpha = 0, beta = 0.25, lr = 0.02;
  bool col_major = false, sync_comm = false;
  std::string input_file = argc > 1 ? argv[1] : "", output_file = "out";
  auto reader = std::unique_ptr<ParallelIO>(new ParallelIO());
  auto grid = std::unique_ptr<Process3DGrid>(new Process3DGrid(world_size, 1, 1, 1));
  auto shared_sparseMat = std::shared_ptr<distblas::core::SpMat<VALUE_TYPE>>(
      new distblas::core::SpMat<VALUE_TYPE>(grid.get()));
  reader.get()->parallel_read_MM<int64_t, int, VALUE_TYPE>(input_file, shared_sparseMat.get(), true);
  auto localARows = shared_sparseMat.get()->gRows;
  shared_sparseMat.get()->batch_size = batch_size;
  shared_sparseMat.get()->initialize_CSR_blocks(true);
  auto dense_mat = std::shared_ptr<DenseMat<INDEX_TYPE, VALUE_TYPE, dimension>>(
      new DenseMat<INDEX_TYPE, VALUE_TYPE, dimension>(grid.get(), localARows));
  std::unique_ptr<distblas::algo::EmbeddingAlgo<INDEX_TYPE, VALUE_TYPE, dimension>> embedding_algo =
      std::unique_ptr<distblas::algo::EmbeddingAlgo<INDEX_TYPE, VALUE_TYPE, dimension>>(
          new distblas::algo::EmbeddingAlgo<INDEX_TYPE, VALUE_TYPE, dimension>(
              shared_sparseMat.get(), shared_sparseMat.get(), shared_sparseMat.get(),
              dense_mat.get(), grid.get(), alpha, beta, 5, -5, col_major, sync_comm));
  embedding_algo.get()->algo_force2_vec_ns(iterations, batch_size, ns, lr);
  reader->parallel_write(output_file + "/embedding.txt", dense_mat.get()->nCoordinates,
                         localARows, dimension, grid.get(), shared_sparseMat.get());
  return embedding_algo->scale(7.0) == 5.0 ? 0 : 1;
}
"""


@pytest.mark.parametrize("dim", [128, 2, 100])
def test_reference_driver_shape_compiles_against_mirror(tmp_path, dim):
    # the reference's dimension is a compile-time constant of any value
    src = tmp_path / "shape.cpp"
    src.write_text(REFERENCE_SHAPED.replace("DIMENSION", str(dim)))
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only",
                        f"-I{ROOT / 'include'}", str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _driver(*args, cwd=None):
    if not DRIVER.exists():
        pytest.skip("driver not built")
    return subprocess.run([str(DRIVER), *args], capture_output=True, text=True, cwd=cwd,
                          timeout=120)


def test_driver_rejects_algorithms_not_built(tmp_path):
    for flag in ("-spgemm", "-sparse_embedding", "-msbfs"):
        r = _driver(flag, "1", "-sbm", "100", cwd=tmp_path)
        assert r.returncode == 2 and "only the force-model embedding" in r.stderr


def test_driver_argument_errors(tmp_path):
    r = _driver("-iter", "1", cwd=tmp_path)
    assert r.returncode == 2 and "-input" in r.stderr
    r = _driver("-sbm", "100", "-dim", "24", cwd=tmp_path)
    assert r.returncode == 2 and "-dim" in r.stderr
    r = _driver("-input", str(tmp_path / "missing.mtx"), cwd=tmp_path)
    assert r.returncode == 1 and "parallel_read_MM" in r.stderr


def test_driver_without_device(tmp_path):
    if _lib.device_count() > 0:
        pytest.skip("a GPU is present")
    r = _driver("-sbm", "200", "-iter", "1", "-batch", "64", "-output", str(tmp_path / "o"),
                cwd=tmp_path)
    assert r.returncode == 1 and "no HIP device" in r.stderr
