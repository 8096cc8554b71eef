// npy_io.h -- particle distributions and snapshots as directories of NumPy .npy files.
//
// The reference reads its initial distribution from one HDF5 file with the datasets /m,
// /x (N x 3), /v (N x 3), /materialId, /u and (INTEGRATE_DENSITY) /rho
// (Miluphpc::distributionFromFile, miluphpc.cpp:157-300) and writes one ts%06d.h5 file per
// output step with /time, /x, /v, /key, /m, /proc, /rho, /p, /e, /sml, /noi, /cs
// (Miluphpc::particles2file, miluphpc.cpp:3451-3760).  HDF5 is not part of this build, so the
// same datasets are stored one per file, <dataset>.npy, in a directory (ts%06d/ for a
// snapshot): the names, shapes and element types are the reference's, and numpy, h5py-based
// converters and the tests read them directly.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace npy {

// element types of the arrays read and written here
enum class Type { f8, i4, i1, u8, i8 };

struct Array {
    Type type = Type::f8;
    std::vector<int64_t> shape;
    std::vector<unsigned char> data;   // little-endian, C order
    int64_t count() const;
};

// Read path into `out`.  Returns an empty string on success, else the reason.
std::string read(const std::string &path, Array &out);
// Write a C-order array of `shape` (1 or 2 dimensions) of `type` from `data`.
std::string write(const std::string &path, Type type, const std::vector<int64_t> &shape, const void *data);

// Element i (row-major) of a numeric array converted to double / int64.
double as_double(const Array &a, int64_t i);
int64_t as_int(const Array &a, int64_t i);

bool exists(const std::string &path);
std::string make_dir(const std::string &path);   // creates path (one level); "" on success

}  // namespace npy
