/*
 * aprfile.cpp -- libaprh_io: the APR file format (APRFile, src/io/APRFile.hpp:28-53,
 * :178-356; APRWriter::FileStructure, src/io/APRWriter.hpp:595-740) for device APRs.
 *
 * Layout (what the reference writes and reads):
 *   /ParticleRepr                 group
 *   /ParticleRepr/<channel>[t]    group per time point ("t", "t1", ...; t = 0 has no number)
 *       attributes: time_steps, name, githash, x_num, y_num, z_num (image dims, u64),
 *                   total_number_particles, level_max, level_min (u64), x_num_<l>, y_num_<l>,
 *                   z_num_<l> (int, l_min <= l < l_max), the pipeline parameters (float)
 *       y_vec (u16), xz_end_vec (u64): LinearAccess, blosc-compressed chunked datasets
 *       <particles name>: particle values, attributes lzxy_order, compress_type,
 *                   quantization_factor, compress_background
 *   /ParticleRepr/<channel>[t]/Tree   the tree's LinearAccess and its particle count
 *
 * Compressed datasets use HDF5 filter 32001 ("blosc").  Its parameter block is the one the
 * published HDF5 blosc filter defines: {filter revision, blosc format version, type size,
 * chunk bytes, compression level, shuffle, compressor code}; the filter itself (below) is
 * this library's own, on the c-blosc library of the system.
 *
 * Some u64 attributes of files the reference wrote carry only 32 significant bits (it writes
 * them from 32-bit integers): readers take the low 32 bits of the image dims and recompute
 * the level geometry from them, as APRFile::read_apr does (APRWriter::read_access_info).
 */
#include <hdf5.h>
#include <blosc.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/aprh.h"
#include "../../include/aprh_io.h"

namespace {

constexpr H5Z_filter_t kBloscFilter = 32001;
constexpr unsigned kBloscFilterRevision = 2;
constexpr hsize_t kChunk = 100000;

struct IoError : std::runtime_error {
    int code;
    IoError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

thread_local std::string g_err;

/* ---------------- the blosc HDF5 filter ---------------- */

/* fills the type size and chunk bytes of the dataset being created */
herr_t blosc_set_local(hid_t dcpl, hid_t type, hid_t) {
    unsigned flags = 0;
    size_t nelem = 8;
    unsigned cd[8] = {0};
    if (H5Pget_filter_by_id2(dcpl, kBloscFilter, &flags, &nelem, cd, 0, nullptr, nullptr) < 0) return -1;
    if (nelem < 4) nelem = 4;
    hsize_t chunk[32];
    const int ndims = H5Pget_chunk(dcpl, 32, chunk);
    if (ndims < 0) return -1;
    const size_t tsize = H5Tget_size(type);
    if (tsize == 0) return -1;
    size_t bytes = tsize;
    for (int i = 0; i < ndims; ++i) bytes *= chunk[i];
    cd[0] = kBloscFilterRevision;
    cd[1] = BLOSC_VERSION_FORMAT;
    cd[2] = (unsigned)tsize;
    cd[3] = (unsigned)bytes;
    return H5Pmodify_filter(dcpl, kBloscFilter, flags, std::max<size_t>(nelem, 7), cd);
}

size_t blosc_filter(unsigned flags, size_t nelem, const unsigned cd[], size_t nbytes, size_t* buf_size, void** buf) {
    if (flags & H5Z_FLAG_REVERSE) {
        size_t out_bytes = 0, cbytes = 0, block = 0;
        blosc_cbuffer_sizes(*buf, &out_bytes, &cbytes, &block);
        void* out = std::malloc(out_bytes ? out_bytes : 1);
        if (!out) return 0;
        const int n = blosc_decompress(*buf, out, out_bytes);
        if (n <= 0) {
            std::free(out);
            return 0;
        }
        std::free(*buf);
        *buf = out;
        *buf_size = out_bytes;
        return (size_t)n;
    }
    const size_t typesize = nelem > 2 ? cd[2] : 1;
    const int clevel = nelem > 4 ? (int)cd[4] : 5;
    const int shuffle = nelem > 5 ? (int)cd[5] : 1;
    const int code = nelem > 6 ? (int)cd[6] : BLOSC_BLOSCLZ;
    const char* name = nullptr;
    if (blosc_compcode_to_compname(code, &name) < 0 || blosc_set_compressor(name) < 0) return 0;
    const size_t cap = nbytes + BLOSC_MAX_OVERHEAD;
    void* out = std::malloc(cap);
    if (!out) return 0;
    const int n = blosc_compress(clevel, shuffle, typesize, nbytes, *buf, out, cap);
    if (n <= 0) {
        std::free(out);
        return 0;
    }
    std::free(*buf);
    *buf = out;
    *buf_size = cap;
    return (size_t)n;
}

void register_blosc() {
    static std::once_flag once;
    std::call_once(once, [] {
        blosc_init();
        H5Z_class2_t cls;
        std::memset(&cls, 0, sizeof(cls));
        cls.version = H5Z_CLASS_T_VERS;
        cls.id = kBloscFilter;
        cls.encoder_present = 1;
        cls.decoder_present = 1;
        cls.name = "blosc";
        cls.can_apply = nullptr;
        cls.set_local = blosc_set_local;
        cls.filter = blosc_filter;
        if (H5Zregister(&cls) < 0) throw IoError(APRH_ERR_HIP, "could not register the blosc HDF5 filter");
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);   /* errors come back as statuses */
    });
}

/* ---------------- small HDF5 helpers ---------------- */

bool link_exists(hid_t loc, const std::string& path) {
    /* every component must exist (H5Lexists only tests the last) */
    size_t pos = 0;
    while (true) {
        pos = path.find('/', pos + 1);
        const std::string p = path.substr(0, pos);
        if (H5Lexists(loc, p.c_str(), H5P_DEFAULT) <= 0) return false;
        if (pos == std::string::npos) return true;
    }
}

template <typename T> hid_t h5type();
template <> hid_t h5type<uint64_t>() { return H5T_NATIVE_UINT64; }
template <> hid_t h5type<uint16_t>() { return H5T_NATIVE_UINT16; }
template <> hid_t h5type<int>() { return H5T_NATIVE_INT; }
template <> hid_t h5type<float>() { return H5T_NATIVE_FLOAT; }

template <typename T>
void write_attr(hid_t loc, const std::string& name, T v) {
    if (H5Aexists(loc, name.c_str()) > 0) H5Adelete(loc, name.c_str());
    hsize_t one = 1;
    const hid_t sp = H5Screate_simple(1, &one, nullptr);
    const hid_t a = H5Acreate2(loc, name.c_str(), h5type<T>(), sp, H5P_DEFAULT, H5P_DEFAULT);
    const herr_t e = (a >= 0) ? H5Awrite(a, h5type<T>(), &v) : -1;
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    if (e < 0) throw IoError(APRH_ERR_HIP, "could not write attribute " + name);
}

void write_attr_string(hid_t loc, const std::string& name, const std::string& v) {
    if (H5Aexists(loc, name.c_str()) > 0) H5Adelete(loc, name.c_str());
    const hid_t t = H5Tcopy(H5T_C_S1);
    H5Tset_size(t, v.size() + 1);
    H5Tset_strpad(t, H5T_STR_NULLTERM);
    const hid_t sp = H5Screate(H5S_SCALAR);
    const hid_t a = H5Acreate2(loc, name.c_str(), t, sp, H5P_DEFAULT, H5P_DEFAULT);
    const herr_t e = (a >= 0) ? H5Awrite(a, t, v.c_str()) : -1;
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    H5Tclose(t);
    if (e < 0) throw IoError(APRH_ERR_HIP, "could not write attribute " + name);
}

/* the attribute as u64 / float / int; false when absent */
bool read_attr_u64(hid_t loc, const char* name, uint64_t* out) {
    if (H5Aexists(loc, name) <= 0) return false;
    const hid_t a = H5Aopen(loc, name, H5P_DEFAULT);
    const herr_t e = H5Aread(a, H5T_NATIVE_UINT64, out);
    H5Aclose(a);
    return e >= 0;
}
bool read_attr_f32(hid_t loc, const char* name, float* out) {
    if (H5Aexists(loc, name) <= 0) return false;
    const hid_t a = H5Aopen(loc, name, H5P_DEFAULT);
    const herr_t e = H5Aread(a, H5T_NATIVE_FLOAT, out);
    H5Aclose(a);
    return e >= 0;
}

template <typename T>
void write_dataset(hid_t loc, const std::string& name, const T* data, uint64_t n, hid_t mem_type, unsigned clevel, unsigned shuffle,
                   unsigned compressor) {
    if (H5Lexists(loc, name.c_str(), H5P_DEFAULT) > 0) {
        /* overwrite an existing dataset of the same size and type (APRFile.hpp:243-262) */
        const hid_t ds = H5Dopen2(loc, name.c_str(), H5P_DEFAULT);
        const hid_t sp = H5Dget_space(ds);
        hsize_t d = 0;
        H5Sget_simple_extent_dims(sp, &d, nullptr);
        const hid_t ft = H5Dget_type(ds);
        const bool same = d == n && H5Tequal(ft, mem_type) > 0;
        H5Tclose(ft);
        herr_t e = -1;
        if (same && n > 0) e = H5Dwrite(ds, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
        H5Sclose(sp);
        H5Dclose(ds);
        if (!same) throw IoError(APRH_ERR_INVALID_ARGUMENT, "dataset " + name + " exists with another size or type");
        if (n > 0 && e < 0) throw IoError(APRH_ERR_HIP, "could not write dataset " + name);
        return;
    }
    hsize_t dims = n, maxd = H5S_UNLIMITED;
    const hid_t sp = H5Screate_simple(1, &dims, &maxd);
    const hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
    hsize_t chunk = std::max<hsize_t>(1, std::min<hsize_t>(n, kChunk));
    H5Pset_chunk(dcpl, 1, &chunk);
    const unsigned cd[7] = {kBloscFilterRevision, BLOSC_VERSION_FORMAT, (unsigned)sizeof(T), (unsigned)(chunk * sizeof(T)), clevel, shuffle,
                            compressor};
    H5Pset_filter(dcpl, kBloscFilter, H5Z_FLAG_OPTIONAL, 7, cd);
    const hid_t ds = H5Dcreate2(loc, name.c_str(), mem_type, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT);
    herr_t e = ds >= 0 ? 0 : -1;
    if (ds >= 0 && n > 0) e = H5Dwrite(ds, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    if (ds >= 0) H5Dclose(ds);
    H5Pclose(dcpl);
    H5Sclose(sp);
    if (e < 0) throw IoError(APRH_ERR_HIP, "could not create dataset " + name);
}

/* size of a 1-D dataset; -1 when absent */
int64_t dataset_size(hid_t loc, const char* name, hid_t* file_type) {
    if (H5Lexists(loc, name, H5P_DEFAULT) <= 0) return -1;
    const hid_t ds = H5Dopen2(loc, name, H5P_DEFAULT);
    if (ds < 0) return -1;
    const hid_t sp = H5Dget_space(ds);
    hsize_t d = 0;
    H5Sget_simple_extent_dims(sp, &d, nullptr);
    if (file_type) *file_type = H5Dget_type(ds);
    H5Sclose(sp);
    H5Dclose(ds);
    return (int64_t)d;
}

void read_dataset(hid_t loc, const char* name, hid_t mem_type, void* out) {
    const hid_t ds = H5Dopen2(loc, name, H5P_DEFAULT);
    if (ds < 0) throw IoError(APRH_ERR_INVALID_ARGUMENT, std::string("no dataset ") + name);
    const herr_t e = H5Dread(ds, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, out);
    H5Dclose(ds);
    if (e < 0) throw IoError(APRH_ERR_HIP, std::string("could not read dataset ") + name + " (compressed with an unsupported filter?)");
}

std::string group_of(uint64_t t, const char* channel) {
    std::string c = channel && *channel ? channel : "t";
    if (t != 0) c += std::to_string(t);
    return "ParticleRepr/" + c;
}

/* GenInfo::init (GenInfo.hpp:33-65) */
struct Geo {
    int64_t org[3] = {0, 0, 0};
    int l_min = 1, l_max = 1;
    std::vector<int64_t> y, x, z;
    void init(int64_t yy, int64_t xx, int64_t zz) {
        org[0] = yy; org[1] = xx; org[2] = zz;
        const int64_t md = std::max(std::max(yy, xx), zz);
        l_max = std::max(1, (int)std::ceil(std::log2((double)md)));
        l_min = 1;
        y.assign(l_max + 1, 0);
        x.assign(l_max + 1, 0);
        z.assign(l_max + 1, 0);
        for (int l = l_min; l <= l_max; ++l) {
            const double cs = std::round(std::pow(2.0, (double)(l_max - l)));
            y[l] = (int64_t)std::ceil((double)yy / cs);
            x[l] = (int64_t)std::ceil((double)xx / cs);
            z[l] = (int64_t)std::ceil((double)zz / cs);
        }
    }
};

}  // namespace

struct aprh_file {
    hid_t fid = -1;
    hid_t root = -1;   /* /ParticleRepr */
    bool writable = false;
};

namespace {

hid_t open_group(aprh_file* f, const std::string& g, bool create) {
    if (link_exists(f->fid, g)) return H5Gopen2(f->fid, g.c_str(), H5P_DEFAULT);
    if (!create) throw IoError(APRH_ERR_INVALID_ARGUMENT, "no time point " + g + " in the file");
    if (!f->writable) throw IoError(APRH_ERR_INVALID_ARGUMENT, "file opened for reading");
    const hid_t id = H5Gcreate2(f->fid, g.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (id < 0) throw IoError(APRH_ERR_HIP, "could not create group " + g);
    return id;
}

struct Group {
    hid_t id;
    explicit Group(hid_t i) : id(i) {}
    ~Group() {
        if (id >= 0) H5Gclose(id);
    }
};

void check_file(aprh_file* f) {
    if (!f || f->fid < 0) throw IoError(APRH_ERR_INVALID_ARGUMENT, "file is not open");
}

void write_structure(aprh_file* f, uint64_t t, const char* channel, const int64_t dims3[3], const uint64_t* xz_end, uint64_t n_xz,
                     const uint16_t* y_vec, uint64_t n_y, const aprh_params* par, const uint64_t* t_xz_end, uint64_t t_n_xz,
                     const uint16_t* t_y_vec, uint64_t t_n_y) {
    check_file(f);
    if (!f->writable) throw IoError(APRH_ERR_INVALID_ARGUMENT, "file opened for reading");
    Geo g;
    g.init(dims3[0], dims3[1], dims3[2]);
    Group grp(open_group(f, group_of(t, channel), true));
    /* APRFile::write_apr (APRFile.hpp:178-204), APRWriter::write_apr_info / write_apr_parameters */
    write_attr<uint64_t>(grp.id, "time_steps", t);
    write_attr_string(grp.id, "name", "no_name");
    write_attr_string(grp.id, "githash", "libaprh");
    write_attr<uint64_t>(grp.id, "x_num", (uint64_t)dims3[1]);
    write_attr<uint64_t>(grp.id, "y_num", (uint64_t)dims3[0]);
    write_attr<uint64_t>(grp.id, "z_num", (uint64_t)dims3[2]);
    write_attr<uint64_t>(grp.id, "total_number_particles", n_y);
    write_attr<uint64_t>(grp.id, "level_max", (uint64_t)g.l_max);
    write_attr<uint64_t>(grp.id, "level_min", (uint64_t)g.l_min);
    for (int l = g.l_min; l < g.l_max; ++l) {
        write_attr<int>(grp.id, "x_num_" + std::to_string(l), (int)g.x[l]);
        write_attr<int>(grp.id, "y_num_" + std::to_string(l), (int)g.y[l]);
        write_attr<int>(grp.id, "z_num_" + std::to_string(l), (int)g.z[l]);
    }
    aprh_params p = par ? *par : aprh_params_default();
    write_attr<float>(grp.id, "lambda", p.lambda);
    write_attr<float>(grp.id, "sigma_th", p.sigma_th);
    write_attr<float>(grp.id, "sigma_th_max", p.sigma_th_max);
    write_attr<float>(grp.id, "I_th", p.Ip_th);
    write_attr<float>(grp.id, "dx", p.dx);
    write_attr<float>(grp.id, "dy", p.dy);
    write_attr<float>(grp.id, "dz", p.dz);
    write_attr<float>(grp.id, "psfx", p.psfx);
    write_attr<float>(grp.id, "psfy", p.psfy);
    write_attr<float>(grp.id, "psfz", p.psfz);
    write_attr<float>(grp.id, "rel_error", p.rel_error);
    write_attr<float>(grp.id, "noise_sd_estimate", 0.0f);
    write_attr<float>(grp.id, "background_intensity_estimate", 0.0f);
    write_attr<float>(grp.id, "grad_th", p.grad_th);
    /* APRWriter::write_linear_access (APRWriter.hpp:357-363): zstd level 4, byte shuffle */
    write_dataset<uint16_t>(grp.id, "y_vec", y_vec, n_y, H5T_NATIVE_UINT16, 4, 1, BLOSC_ZSTD);
    write_dataset<uint64_t>(grp.id, "xz_end_vec", xz_end, n_xz, H5T_NATIVE_UINT64, 4, 1, BLOSC_ZSTD);
    if (t_xz_end) {
        Group tree(open_group(f, group_of(t, channel) + "/Tree", true));
        write_attr<uint64_t>(tree.id, "total_number_particles", t_n_y);
        write_dataset<uint16_t>(tree.id, "y_vec", t_y_vec, t_n_y, H5T_NATIVE_UINT16, 4, 1, BLOSC_ZSTD);
        write_dataset<uint64_t>(tree.id, "xz_end_vec", t_xz_end, t_n_xz, H5T_NATIVE_UINT64, 4, 1, BLOSC_ZSTD);
    }
}

/* dims (low 32 bits of the stored u64, see the file comment), total, dataset sizes */
void read_info(aprh_file* f, uint64_t t, const char* channel, int64_t dims3[3], uint64_t* total, int64_t* n_xz, int64_t* n_y) {
    check_file(f);
    Group grp(open_group(f, group_of(t, channel), false));
    const char* names[3] = {"y_num", "x_num", "z_num"};
    for (int i = 0; i < 3; ++i) {
        uint64_t v = 0;
        if (!read_attr_u64(grp.id, names[i], &v)) throw IoError(APRH_ERR_INVALID_ARGUMENT, std::string("no attribute ") + names[i]);
        dims3[i] = (int64_t)(v & 0xFFFFFFFFull);
        if (dims3[i] <= 0) throw IoError(APRH_ERR_BAD_DIMS, "image dims in the file are not positive");
    }
    uint64_t tot = 0;
    read_attr_u64(grp.id, "total_number_particles", &tot);
    if (total) *total = tot;
    const int64_t nx = dataset_size(grp.id, "xz_end_vec", nullptr), ny = dataset_size(grp.id, "y_vec", nullptr);
    if (nx < 0 || ny < 0)
        throw IoError(APRH_ERR_UNSUPPORTED, "no linear access (y_vec / xz_end_vec) at this time point (random-access files are not read)");
    if (n_xz) *n_xz = nx;
    if (n_y) *n_y = ny;
}

}  // namespace

extern "C" {

const char* aprh_io_last_error(void) { return g_err.c_str(); }

#define IO_TRY try {
#define IO_CATCH                            \
    }                                       \
    catch (IoError & e) {                   \
        g_err = e.what();                   \
        return e.code;                      \
    }                                       \
    catch (std::exception & e) {            \
        g_err = e.what();                   \
        return APRH_ERR_HIP;                \
    }                                       \
    return APRH_OK;

int aprh_file_open(const char* path, const char* mode, aprh_file** out) {
    IO_TRY
    if (!path || !mode || !out) throw IoError(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    register_blosc();
    const std::string m = mode;
    aprh_file* f = new aprh_file();
    if (m == "READ") {
        f->fid = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    } else if (m == "WRITE") {
        f->fid = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
        f->writable = true;
    } else if (m == "READWRITE") {
        f->fid = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
        f->writable = true;
    } else {
        delete f;
        throw IoError(APRH_ERR_INVALID_ARGUMENT, "mode must be READ, WRITE or READWRITE");
    }
    if (f->fid < 0) {
        delete f;
        throw IoError(APRH_ERR_INVALID_ARGUMENT, std::string("could not open ") + path);
    }
    if (link_exists(f->fid, "ParticleRepr")) f->root = H5Gopen2(f->fid, "ParticleRepr", H5P_DEFAULT);
    else if (f->writable) f->root = H5Gcreate2(f->fid, "ParticleRepr", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (f->root < 0) {
        H5Fclose(f->fid);
        delete f;
        throw IoError(APRH_ERR_INVALID_ARGUMENT, "not an APR file (no ParticleRepr group)");
    }
    *out = f;
    IO_CATCH
}

int aprh_file_close(aprh_file* f) {
    IO_TRY
    if (!f) return APRH_OK;
    if (f->root >= 0) H5Gclose(f->root);
    if (f->fid >= 0) H5Fclose(f->fid);
    delete f;
    IO_CATCH
}

int aprh_file_read_info(aprh_file* f, uint64_t t, const char* channel, int64_t* dims3, uint64_t* total_particles, uint64_t* xz_end_size) {
    IO_TRY
    if (!dims3) throw IoError(APRH_ERR_INVALID_ARGUMENT, "NULL dims");
    int64_t nx = 0;
    read_info(f, t, channel, dims3, total_particles, &nx, nullptr);
    if (xz_end_size) *xz_end_size = (uint64_t)nx;
    IO_CATCH
}

int aprh_file_read_structure(aprh_file* f, uint64_t t, const char* channel, uint64_t* xz_end_vec, uint16_t* y_vec) {
    IO_TRY
    int64_t dims[3], nx = 0, ny = 0;
    read_info(f, t, channel, dims, nullptr, &nx, &ny);
    Group grp(open_group(f, group_of(t, channel), false));
    if (xz_end_vec && nx > 0) read_dataset(grp.id, "xz_end_vec", H5T_NATIVE_UINT64, xz_end_vec);
    if (y_vec && ny > 0) read_dataset(grp.id, "y_vec", H5T_NATIVE_UINT16, y_vec);
    IO_CATCH
}

int aprh_file_write_structure(aprh_file* f, uint64_t t, const char* channel, const int64_t* dims3, const uint64_t* xz_end_vec,
                              uint64_t xz_end_size, const uint16_t* y_vec, uint64_t total_particles, const aprh_params* par) {
    IO_TRY
    if (!dims3 || !xz_end_vec || (total_particles && !y_vec)) throw IoError(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    write_structure(f, t, channel, dims3, xz_end_vec, xz_end_size, y_vec, total_particles, par, nullptr, 0, nullptr, 0);
    IO_CATCH
}

int aprh_file_write_apr(aprh_context* ctx, aprh_file* f, aprh_apr* apr, uint64_t t, const char* channel, int with_tree,
                        const aprh_params* par) {
    IO_TRY
    if (!ctx || !apr) throw IoError(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    aprh_apr_info info;
    int st = aprh_apr_info_get(apr, &info);
    if (st) throw IoError(st, aprh_last_error());
    std::vector<uint64_t> lxz(info.level_xz_size), xze(info.xz_end_size);
    std::vector<uint16_t> yv(std::max<uint64_t>(info.total_number_particles, 1));
    st = aprh_apr_copy_structure(apr, lxz.data(), xze.data(), yv.data());
    if (st) throw IoError(st, aprh_last_error());
    std::vector<uint64_t> txze;
    std::vector<uint16_t> tyv;
    uint64_t tn = 0;
    if (with_tree) {
        aprh_tree_info ti;
        st = aprh_apr_tree_info(ctx, apr, &ti);
        if (st) throw IoError(st, aprh_last_error());
        std::vector<uint64_t> tlxz(ti.level_xz_size);
        txze.resize(ti.xz_end_size);
        tyv.resize(std::max<uint64_t>(ti.total_number_tree_particles, 1));
        tn = ti.total_number_tree_particles;
        st = aprh_apr_copy_tree(ctx, apr, tlxz.data(), txze.data(), tyv.data());
        if (st) throw IoError(st, aprh_last_error());
    }
    write_structure(f, t, channel, info.org_dims, xze.data(), xze.size(), yv.data(), info.total_number_particles, par,
                    with_tree ? txze.data() : nullptr, txze.size(), tyv.data(), tn);
    IO_CATCH
}

int aprh_file_read_apr(aprh_context* ctx, aprh_file* f, uint64_t t, const char* channel, aprh_apr** out) {
    IO_TRY
    if (!ctx || !out) throw IoError(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    int64_t dims[3], nx = 0, ny = 0;
    read_info(f, t, channel, dims, nullptr, &nx, &ny);
    std::vector<uint64_t> xze((size_t)std::max<int64_t>(nx, 1));
    std::vector<uint16_t> yv((size_t)std::max<int64_t>(ny, 1));
    Group grp(open_group(f, group_of(t, channel), false));
    if (nx > 0) read_dataset(grp.id, "xz_end_vec", H5T_NATIVE_UINT64, xze.data());
    if (ny > 0) read_dataset(grp.id, "y_vec", H5T_NATIVE_UINT16, yv.data());
    const int st = aprh_apr_from_structure(ctx, dims[0], dims[1], dims[2], xze.data(), (uint64_t)nx, yv.data(), (uint64_t)ny, out);
    if (st) throw IoError(st, aprh_last_error());
    IO_CATCH
}

int aprh_file_write_particles(aprh_context* ctx, aprh_file* f, const char* name, const void* parts, int dtype, uint64_t n, int apr_or_tree,
                              uint64_t t, const char* channel) {
    IO_TRY
    check_file(f);
    if (!name || (n && !parts)) throw IoError(APRH_ERR_INVALID_ARGUMENT, "NULL argument");
    if (dtype != APRH_DTYPE_U16 && dtype != APRH_DTYPE_F32) throw IoError(APRH_ERR_INVALID_ARGUMENT, "dtype must be u16 or f32");
    const size_t es = dtype == APRH_DTYPE_U16 ? 2 : 4;
    std::vector<uint8_t> host(std::max<size_t>(n * es, 1));
    if (n && ctx) {
        const int st = aprh_memcpy(ctx, host.data(), parts, n * es);
        if (st) throw IoError(st, aprh_last_error());
    } else if (n) {
        std::memcpy(host.data(), parts, n * es);   /* no context: host memory */
    }
    const std::string g = group_of(t, channel) + (apr_or_tree ? "" : "/Tree");
    Group grp(open_group(f, g, true));
    /* APRFile::write_particles (APRFile.hpp:211-270): uncompressed values (compress_type 0) */
    write_attr<int>(grp.id, "lzxy_order", 1);
    write_attr<int>(grp.id, "compress_type", 0);
    write_attr<float>(grp.id, "quantization_factor", 0.5f);
    write_attr<float>(grp.id, "compress_background", 1.0f);
    if (dtype == APRH_DTYPE_U16)
        write_dataset<uint16_t>(grp.id, name, (const uint16_t*)host.data(), n, H5T_NATIVE_UINT16, 2, 1, BLOSC_ZSTD);
    else
        write_dataset<float>(grp.id, name, (const float*)host.data(), n, H5T_NATIVE_FLOAT, 2, 1, BLOSC_ZSTD);
    IO_CATCH
}

int aprh_file_read_particles(aprh_context* ctx, aprh_file* f, const char* name, int apr_or_tree, uint64_t t, const char* channel, void* out,
                             uint64_t capacity, int* dtype, uint64_t* n) {
    IO_TRY
    check_file(f);
    if (!name) throw IoError(APRH_ERR_INVALID_ARGUMENT, "NULL name");
    const std::string g = group_of(t, channel) + (apr_or_tree ? "" : "/Tree");
    Group grp(open_group(f, g, false));
    hid_t ft = -1;
    const int64_t sz = dataset_size(grp.id, name, &ft);
    if (sz < 0) throw IoError(APRH_ERR_INVALID_ARGUMENT, std::string("no particle dataset ") + name);
    const H5T_class_t cls = H5Tget_class(ft);
    const size_t fsz = H5Tget_size(ft);
    H5Tclose(ft);
    int dt;
    if (cls == H5T_INTEGER && fsz == 2) dt = APRH_DTYPE_U16;
    else if (cls == H5T_FLOAT && fsz == 4) dt = APRH_DTYPE_F32;
    else throw IoError(APRH_ERR_UNSUPPORTED, "particle datasets of u16 or f32 only");
    int ct = 0;
    if (H5Aexists(grp.id, "compress_type") > 0) {
        const hid_t a = H5Aopen(grp.id, "compress_type", H5P_DEFAULT);
        H5Aread(a, H5T_NATIVE_INT, &ct);
        H5Aclose(a);
    }
    if (ct != 0) throw IoError(APRH_ERR_UNSUPPORTED, "compressed particle values (APRCompress) are not read");
    if (dtype) *dtype = dt;
    if (n) *n = (uint64_t)sz;
    if (!out) return APRH_OK;
    if (capacity < (uint64_t)sz) throw IoError(APRH_ERR_INVALID_ARGUMENT, "particle buffer too small");
    const size_t es = dt == APRH_DTYPE_U16 ? 2 : 4;
    const hid_t mt = dt == APRH_DTYPE_U16 ? H5T_NATIVE_UINT16 : H5T_NATIVE_FLOAT;
    if (!ctx) {
        /* no context: host memory */
        if (sz) read_dataset(grp.id, name, mt, out);
        return APRH_OK;
    }
    std::vector<uint8_t> host(std::max<size_t>((size_t)sz * es, 1));
    if (sz) read_dataset(grp.id, name, mt, host.data());
    if (sz) {
        const int st = aprh_memcpy(ctx, out, host.data(), (size_t)sz * es);
        if (st) throw IoError(st, aprh_last_error());
    }
    IO_CATCH
}

}  // extern "C"
