// SimulationFile (HDF5) input and output for the MPD-RDME path; declared in
// include/lmx_io.h.  Host C++ over the HDF5 C and high-level APIs.  The file
// layout (group and dataset names, types, chunking, gzip level, attribute
// names) is the reference's file format, so that files move between the two
// (src/io/SimulationFile.cpp, src/io/SimulationFile_create.cpp,
// docs/HDF5FileFormat.text); the code is this project's own.
#include "lmx_io.h"

#include <hdf5.h>
#include <hdf5_hl.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace {

thread_local std::string g_error;

int fail(int code, const std::string &msg)
{
    g_error = msg;
    return code;
}

constexpr uint32_t kCurrentVersion = 4;   // SimulationFile::CURRENT_VERSION
constexpr uint32_t kMinVersion = 2;       // SimulationFile::MIN_VERSION
constexpr hsize_t kRateCols = LMXIO_MAX_RATE_CONSTANTS;
constexpr hsize_t kNoiseCols = LMXIO_NOISE_COLS;
constexpr hsize_t kGzipChunk = 16;        // TUNE_LATTICE_GZIP_CHUNK_SIZE default (src/cmake/options.cmake:84)
constexpr unsigned kGzipLevel = 1;        // TUNE_LATTICE_GZIP_COMPRESSION_LEVEL default (:85)
constexpr hsize_t kSeriesChunk = 100;     // rows per chunk of the extendible time series

struct Quiet {
    // HDF5 prints its error stack by default; the status codes carry it here.
    Quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
};
const Quiet g_quiet;

std::string replicate_name(uint32_t replicate)
{
    char buf[16];
    std::snprintf(buf, sizeof(buf), "%07u", replicate);
    return std::string("/Simulations/") + buf;
}

bool link_exists(hid_t loc, const std::string &path)
{
    // Each component in turn, so that a missing parent is an answer, not an error.
    size_t pos = 0;
    while (true) {
        pos = path.find('/', pos + 1);
        const std::string part = path.substr(0, pos);
        if (!part.empty() && part != "/" && H5Lexists(loc, part.c_str(), H5P_DEFAULT) <= 0) return false;
        if (pos == std::string::npos) return true;
    }
}

bool attr_exists(hid_t loc, const char *obj, const char *name)
{
    return H5Aexists_by_name(loc, obj, name, H5P_DEFAULT) > 0;
}

struct Ds {
    std::vector<hsize_t> dims;
    size_t size = 0;
    H5T_class_t cls = H5T_NO_CLASS;
    size_t elements() const
    {
        size_t n = 1;
        for (hsize_t d : dims) n *= (size_t)d;
        return n;
    }
};

int dataset_info(hid_t loc, const std::string &path, Ds &d)
{
    int nd = 0;
    if (!link_exists(loc, path) || H5LTget_dataset_ndims(loc, path.c_str(), &nd) < 0)
        return fail(LMXIO_ERR_FORMAT, "missing dataset " + path);
    d.dims.assign((size_t)nd, 0);
    if (H5LTget_dataset_info(loc, path.c_str(), d.dims.data(), &d.cls, &d.size) < 0)
        return fail(LMXIO_ERR_FORMAT, "unreadable dataset " + path);
    return LMXIO_OK;
}

bool dims_are(const Ds &d, std::initializer_list<hsize_t> want, size_t size)
{
    if (d.dims.size() != want.size() || d.size != size) return false;
    size_t k = 0;
    for (hsize_t w : want)
        if (d.dims[k++] != w) return false;
    return true;
}

int make_group(hid_t loc, const char *path)
{
    if (H5Lexists(loc, path, H5P_DEFAULT) > 0) return LMXIO_OK;
    const hid_t g = H5Gcreate2(loc, path, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (g < 0) return fail(LMXIO_ERR_IO, std::string("cannot create group ") + path);
    H5Gclose(g);
    return LMXIO_OK;
}

// An empty extendible dataset: rank 1 (cols == 0) or rank 2 with cols columns.
int create_series(hid_t group, const char *name, hid_t file_type, hsize_t cols)
{
    const int rank = cols ? 2 : 1;
    hsize_t dims[2] = {0, cols}, maxd[2] = {H5S_UNLIMITED, cols}, chunk[2] = {kSeriesChunk, cols};
    const hid_t sp = H5Screate_simple(rank, dims, maxd);
    const hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
    H5Pset_chunk(pl, rank, chunk);
    const hid_t ds = H5Dcreate2(group, name, file_type, sp, H5P_DEFAULT, pl, H5P_DEFAULT);
    H5Pclose(pl);
    H5Sclose(sp);
    if (ds < 0) return fail(LMXIO_ERR_IO, std::string("cannot create ") + name);
    H5Dclose(ds);
    return LMXIO_OK;
}

// Appends rows to an extendible dataset; *first receives the first new row.
int append_rows(hid_t group, const char *name, hid_t mem_type, hsize_t cols, hsize_t rows, const void *data,
                hsize_t *first = nullptr)
{
    const int rank = cols ? 2 : 1;
    const hid_t ds = H5Dopen2(group, name, H5P_DEFAULT);
    if (ds < 0) return fail(LMXIO_ERR_FORMAT, std::string("missing dataset ") + name);
    hsize_t dims[2] = {0, 0};
    hid_t sp = H5Dget_space(ds);
    const int got = H5Sget_simple_extent_ndims(sp);
    H5Sget_simple_extent_dims(sp, dims, nullptr);
    H5Sclose(sp);
    if (got != rank || (rank == 2 && dims[1] != cols)) {
        H5Dclose(ds);
        return fail(LMXIO_ERR_FORMAT, std::string(name) + ": shape differs from the data appended");
    }
    const hsize_t start0 = dims[0];
    if (first) *first = start0;
    dims[0] += rows;
    herr_t st = H5Dset_extent(ds, dims);
    if (st >= 0 && rows) {
        hsize_t mdims[2] = {rows, cols}, start[2] = {start0, 0};
        const hid_t ms = H5Screate_simple(rank, mdims, nullptr);
        sp = H5Dget_space(ds);
        st = H5Sselect_hyperslab(sp, H5S_SELECT_SET, start, nullptr, mdims, nullptr);
        if (st >= 0) st = H5Dwrite(ds, mem_type, ms, sp, H5P_DEFAULT, data);
        H5Sclose(sp);
        H5Sclose(ms);
    }
    H5Dclose(ds);
    if (st < 0) return fail(LMXIO_ERR_IO, std::string("cannot append to ") + name);
    return LMXIO_OK;
}

// A gzip-compressed lattice-shaped dataset (SimulationFile.cpp:760-800):
// chunks of min(16, n) along x, y and z and the whole particle dimension.
int write_frame(hid_t group, const char *name, int rank, const hsize_t *dims, hid_t file_type, hid_t mem_type,
                const void *data)
{
    hsize_t chunk[4];
    for (int k = 0; k < rank; k++) chunk[k] = k < 3 ? std::min(kGzipChunk, dims[k]) : dims[k];
    const hid_t sp = H5Screate_simple(rank, dims, nullptr);
    const hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
    H5Pset_deflate(pl, kGzipLevel);
    H5Pset_chunk(pl, rank, chunk);
    const hid_t ds = H5Dcreate2(group, name, file_type, sp, H5P_DEFAULT, pl, H5P_DEFAULT);
    const herr_t st = ds < 0 ? -1 : H5Dwrite(ds, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    if (ds >= 0) H5Dclose(ds);
    H5Pclose(pl);
    H5Sclose(sp);
    if (st < 0) return fail(LMXIO_ERR_IO, std::string("cannot write ") + name);
    return LMXIO_OK;
}

int lattice_types(uint32_t bytesPerParticle, hid_t *file_type, hid_t *mem_type)
{
    if (bytesPerParticle == 1) {
        *file_type = H5T_STD_U8LE;
        *mem_type = H5T_NATIVE_UINT8;
    } else if (bytesPerParticle == 4) {
        *file_type = H5T_STD_U32LE;
        *mem_type = H5T_NATIVE_UINT32;
    } else {
        return fail(LMXIO_ERR_ARG, "bytes per particle must be 1 or 4");
    }
    return LMXIO_OK;
}

}  // namespace

struct lmxio_file {
    hid_t file = -1;
    uint32_t version = 0;
    bool writable = false;
};

namespace {

// The model's species count, which sizes a new replicate's SpeciesCounts
// (SimulationFile::loadModel, SimulationFile.cpp:338-359).
int model_species(lmxio_file *f, uint32_t *nS)
{
    const char *obj = f->version <= 3 ? "/Model" : "/Model/Reaction";
    if (!link_exists(f->file, obj) || !attr_exists(f->file, obj, "numberSpecies") ||
        H5LTget_attribute_uint(f->file, obj, "numberSpecies", nS) < 0)
        return fail(LMXIO_ERR_FORMAT, "no model has been defined");
    return LMXIO_OK;
}

// Opens /Simulations/%07d, creating it with empty SpeciesCounts [0][nS] and
// SpeciesCountTimes when it does not exist (openReplicateHandles and
// createReplicateHandles, SimulationFile.cpp:2138-2230).
int open_replicate(lmxio_file *f, uint32_t replicate, bool create, hid_t *group)
{
    const std::string rep = replicate_name(replicate);
    if (H5Lexists(f->file, rep.c_str(), H5P_DEFAULT) > 0) {
        *group = H5Gopen2(f->file, rep.c_str(), H5P_DEFAULT);
        return *group < 0 ? fail(LMXIO_ERR_IO, "cannot open " + rep) : LMXIO_OK;
    }
    if (!create) return fail(LMXIO_ERR_FORMAT, "replicate " + std::to_string(replicate) + " does not exist");
    if (!f->writable) return fail(LMXIO_ERR_ARG, "file opened read-only");
    uint32_t nS = 0;
    int rc = model_species(f, &nS);
    if (rc) return rc;
    *group = H5Gcreate2(f->file, rep.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (*group < 0) return fail(LMXIO_ERR_IO, "cannot create " + rep);
    rc = create_series(*group, "SpeciesCounts", H5T_STD_I32LE, nS);
    if (!rc) rc = create_series(*group, "SpeciesCountTimes", H5T_IEEE_F64LE, 0);
    if (rc) H5Gclose(*group);
    return rc;
}

// Appends one frame to group_name (created with its times dataset on first
// use) and its time to times_name (appendLattice, SimulationFile.cpp:1166-1265;
// appendSites, :1347-1444).
int append_frame(lmxio_file *f, uint32_t replicate, const char *group_name, const char *times_name, double time,
                 int rank, const hsize_t *dims, hid_t file_type, hid_t mem_type, const void *data)
{
    if (!f->writable) return fail(LMXIO_ERR_ARG, "file opened read-only");
    hid_t g;
    int rc = open_replicate(f, replicate, true, &g);
    if (rc) return rc;
    if (H5Lexists(g, group_name, H5P_DEFAULT) <= 0) {
        rc = make_group(g, group_name);
        if (!rc) rc = create_series(g, times_name, H5T_IEEE_F64LE, 0);
    }
    hsize_t index = 0;
    if (!rc) rc = append_rows(g, times_name, H5T_NATIVE_DOUBLE, 0, 1, &time, &index);
    if (!rc) {
        const hid_t fg = H5Gopen2(g, group_name, H5P_DEFAULT);
        char name[16];
        std::snprintf(name, sizeof(name), "%010llu", (unsigned long long)index);
        rc = fg < 0 ? fail(LMXIO_ERR_IO, std::string("cannot open ") + group_name)
                    : write_frame(fg, name, rank, dims, file_type, mem_type, data);
        if (fg >= 0) H5Gclose(fg);
    }
    H5Gclose(g);
    return rc;
}

bool valid_file(const char *path)
{
    // isValidFile (SimulationFile.cpp:225-252): the "LMH5" magic, then HDF5.
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return false;
    char magic[4] = {0, 0, 0, 0};
    const size_t got = std::fread(magic, 1, 4, fp);
    std::fclose(fp);
    if (got != 4 || std::memcmp(magic, "LMH5", 4) != 0) return false;
    return H5Fis_hdf5(path) > 0;
}

struct ParamCtx {
    uint32_t version;
    std::map<std::string, std::string> values;
};

// parseParameter (SimulationFile.cpp:260-306): version 2 files hold doubles,
// later ones fixed-length strings; other attributes are skipped.
herr_t param_cb(hid_t loc, const char *name, const H5A_info_t *, void *op)
{
    ParamCtx *ctx = static_cast<ParamCtx *>(op);
    const hid_t a = H5Aopen(loc, name, H5P_DEFAULT);
    if (a < 0) return -1;
    const hid_t t = H5Aget_type(a);
    const H5T_class_t cls = H5Tget_class(t);
    const size_t n = (size_t)H5Aget_storage_size(a);
    herr_t st = 0;
    if (ctx->version == 2 && cls == H5T_FLOAT && n == sizeof(double)) {
        double v = 0.0;
        st = H5Aread(a, H5T_NATIVE_DOUBLE, &v);
        char buf[33];
        std::snprintf(buf, sizeof(buf), "%g", v);
        if (st >= 0) ctx->values[name] = buf;
    } else if (ctx->version >= 3 && cls == H5T_STRING) {
        if (H5Tis_variable_str(t) > 0) {
            char *s = nullptr;
            const hid_t mt = H5Tcopy(H5T_C_S1);
            H5Tset_size(mt, H5T_VARIABLE);
            st = H5Aread(a, mt, &s);
            if (st >= 0) {
                ctx->values[name] = s ? s : "";
                if (s) H5free_memory(s);
            }
            H5Tclose(mt);
        } else {
            std::vector<char> buf(n + 1, 0);
            if (n) {
                const hid_t mt = H5Tcopy(H5T_C_S1);
                H5Tset_size(mt, n);
                st = H5Aread(a, mt, buf.data());
                H5Tclose(mt);
            }
            if (st >= 0) ctx->values[name] = std::string(buf.data(), strnlen(buf.data(), n));
        }
    }
    H5Tclose(t);
    H5Aclose(a);
    return st < 0 ? -1 : 0;
}

}  // namespace

extern "C" {

const char *lmxio_last_error(void) { return g_error.c_str(); }

int lmxio_create(const char *path, int initializeModel, uint32_t numberSpecies)
{
    if (!path) return fail(LMXIO_ERR_ARG, "null path");
    struct stat st;
    if (::stat(path, &st) == 0)
        return fail(LMXIO_ERR_IO, std::string("a file with the specified filename already exists: ") + path);
    const hid_t cp = H5Pcreate(H5P_FILE_CREATE);
    H5Pset_userblock(cp, 512);
    const hid_t f = H5Fcreate(path, H5F_ACC_EXCL, cp, H5P_DEFAULT);
    H5Pclose(cp);
    if (f < 0) return fail(LMXIO_ERR_IO, std::string("cannot create ") + path);
    const unsigned int version = kCurrentVersion;
    int rc = H5LTset_attribute_uint(f, "/", "formatVersion", &version, 1) < 0
                 ? fail(LMXIO_ERR_IO, "cannot write formatVersion") : LMXIO_OK;
    for (const char *g : {"/Parameters", "/Model", "/Simulations"})
        if (!rc) rc = make_group(f, g);
    if (!rc && initializeModel) {
        rc = make_group(f, "/Model/Reaction");
        if (!rc && H5LTset_attribute_uint(f, "/Model/Reaction", "numberSpecies", &numberSpecies, 1) < 0)
            rc = fail(LMXIO_ERR_IO, "cannot write numberSpecies");
    }
    if (H5Fclose(f) < 0 && !rc) rc = fail(LMXIO_ERR_IO, "cannot close the new file");
    if (rc) return rc;
    // The file magic, at the start of the user block.
    FILE *fp = std::fopen(path, "rb+");
    if (!fp) return fail(LMXIO_ERR_IO, std::string("cannot reopen ") + path);
    const bool ok = std::fwrite("LMH5", 1, 4, fp) == 4;
    if (std::fclose(fp) != 0 || !ok) return fail(LMXIO_ERR_IO, "cannot write the file magic");
    return LMXIO_OK;
}

int lmxio_open(const char *path, int writable, lmxio_file **out)
{
    if (!path || !out) return fail(LMXIO_ERR_ARG, "null argument");
    *out = nullptr;
    struct stat st;
    if (::stat(path, &st) != 0) return fail(LMXIO_ERR_IO, std::string("the specified file did not exist: ") + path);
    if (!S_ISREG(st.st_mode)) return fail(LMXIO_ERR_IO, std::string("the specified file was not a regular file: ") + path);
    if (!valid_file(path)) return fail(LMXIO_ERR_IO, std::string("the specified file was not of the correct format: ") + path);
    const hid_t f = H5Fopen(path, writable ? H5F_ACC_RDWR : H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail(LMXIO_ERR_IO, std::string("cannot open ") + path);
    hsize_t ub = 0;
    const hid_t cp = H5Fget_create_plist(f);
    H5Pget_userblock(cp, &ub);
    H5Pclose(cp);
    unsigned int version = 0;
    int rc = LMXIO_OK;
    if (ub < 4) rc = fail(LMXIO_ERR_FORMAT, "the file's user block is too small for the magic");
    else if (!attr_exists(f, "/", "formatVersion") || H5LTget_attribute_uint(f, "/", "formatVersion", &version) < 0)
        rc = fail(LMXIO_ERR_FORMAT, "missing formatVersion");
    else if (version < kMinVersion || version > kCurrentVersion)
        rc = fail(LMXIO_ERR_FORMAT, "unsupported format version " + std::to_string(version));
    for (const char *g : {"/Parameters", "/Model", "/Simulations"})
        if (!rc && H5Lexists(f, g, H5P_DEFAULT) <= 0) rc = fail(LMXIO_ERR_FORMAT, std::string("missing group ") + g);
    if (rc) {
        H5Fclose(f);
        return rc;
    }
    lmxio_file *h = new lmxio_file();
    h->file = f;
    h->version = version;
    h->writable = writable != 0;
    *out = h;
    return LMXIO_OK;
}

int lmxio_close(lmxio_file *f)
{
    if (!f) return LMXIO_OK;
    const herr_t st = f->file >= 0 ? H5Fclose(f->file) : 0;
    delete f;
    return st < 0 ? fail(LMXIO_ERR_IO, "close failed") : LMXIO_OK;
}

int lmxio_format_version(const lmxio_file *f, uint32_t *version)
{
    if (!f || !version) return fail(LMXIO_ERR_ARG, "null argument");
    *version = f->version;
    return LMXIO_OK;
}

int lmxio_flush(lmxio_file *f)
{
    if (!f) return fail(LMXIO_ERR_ARG, "null argument");
    return H5Fflush(f->file, H5F_SCOPE_GLOBAL) < 0 ? fail(LMXIO_ERR_IO, "flush failed") : LMXIO_OK;
}

int lmxio_get_parameters(lmxio_file *f, char *buf, size_t cap, size_t *needed)
{
    if (!f || !needed) return fail(LMXIO_ERR_ARG, "null argument");
    ParamCtx ctx{f->version, {}};
    const hid_t g = H5Gopen2(f->file, "/Parameters", H5P_DEFAULT);
    if (g < 0) return fail(LMXIO_ERR_FORMAT, "missing /Parameters");
    hsize_t idx = 0;
    const herr_t st = H5Aiterate2(g, H5_INDEX_NAME, H5_ITER_INC, &idx, param_cb, &ctx);
    H5Gclose(g);
    if (st < 0) return fail(LMXIO_ERR_FORMAT, "cannot read /Parameters");
    std::string out;
    for (const auto &kv : ctx.values) out += kv.first + "=" + kv.second + "\n";
    *needed = out.size() + 1;
    if (buf) {
        if (cap < *needed) return fail(LMXIO_ERR_ARG, "parameter buffer too small");
        std::memcpy(buf, out.c_str(), *needed);
    }
    return LMXIO_OK;
}

int lmxio_set_parameter(lmxio_file *f, const char *key, const char *value)
{
    if (!f || !key || !value || !*key || std::strchr(key, '=') || std::strchr(key, '\n') || std::strchr(value, '\n'))
        return fail(LMXIO_ERR_ARG, "invalid parameter key or value");
    if (!f->writable) return fail(LMXIO_ERR_ARG, "file opened read-only");
    if (attr_exists(f->file, "/Parameters", key)) H5Adelete_by_name(f->file, "/Parameters", key, H5P_DEFAULT);
    herr_t st;
    if (f->version <= 2) {
        const double d = std::atof(value);
        st = H5LTset_attribute_double(f->file, "/Parameters", key, &d, 1);
    } else {
        st = H5LTset_attribute_string(f->file, "/Parameters", key, value);
    }
    return st < 0 ? fail(LMXIO_ERR_IO, std::string("cannot write parameter ") + key) : LMXIO_OK;
}

int lmxio_get_reaction_info(lmxio_file *f, lmxio_reaction_info *info)
{
    if (!f || !info) return fail(LMXIO_ERR_ARG, "null argument");
    std::memset(info, 0, sizeof(*info));
    const hid_t h = f->file;
    const char *g = "/Model/Reaction";
    if (!link_exists(h, g)) return LMXIO_OK;
    info->present = 1;
    if (H5LTget_attribute_uint(h, g, "numberSpecies", &info->numberSpecies) < 0)
        return fail(LMXIO_ERR_FORMAT, "/Model/Reaction: missing numberSpecies");
    if (attr_exists(h, g, "numberReactions")) {
        info->hasTables = 1;
        if (H5LTget_attribute_uint(h, g, "numberReactions", &info->numberReactions) < 0)
            return fail(LMXIO_ERR_FORMAT, "/Model/Reaction: unreadable numberReactions");
        info->hasRateNoise = info->numberReactions && link_exists(h, "/Model/Reaction/ReactionRateNoise");
    }
    return LMXIO_OK;
}

int lmxio_get_reaction_model(lmxio_file *f, uint32_t *initialCounts, uint32_t *types, double *K, double *noise,
                             int32_t *S, uint32_t *D)
{
    lmxio_reaction_info info;
    int rc = lmxio_get_reaction_info(f, &info);
    if (rc) return rc;
    if (!info.hasTables) return LMXIO_OK;  // only the number of species is defined
    const hid_t h = f->file;
    const uint32_t nS = info.numberSpecies, R = info.numberReactions;
    Ds d;
    if ((rc = dataset_info(h, "/Model/Reaction/InitialSpeciesCounts", d))) return rc;
    if (d.dims.size() != 1 || d.dims[0] != nS || d.size != sizeof(uint32_t))
        return fail(LMXIO_ERR_FORMAT, "/Model/Reaction/InitialSpeciesCounts: invalid dataset dimensions");
    if (initialCounts && H5LTread_dataset(h, "/Model/Reaction/InitialSpeciesCounts", H5T_NATIVE_UINT32, initialCounts) < 0)
        return fail(LMXIO_ERR_IO, "cannot read InitialSpeciesCounts");
    if (!R) return LMXIO_OK;
    struct Check { const char *name; std::initializer_list<hsize_t> dims; size_t size; };
    const Check checks[] = {
        {"/Model/Reaction/ReactionTypes", {R}, sizeof(uint32_t)},
        {"/Model/Reaction/ReactionRateConstants", {R, kRateCols}, sizeof(double)},
        {"/Model/Reaction/StoichiometricMatrix", {nS, R}, sizeof(int32_t)},
        {"/Model/Reaction/DependencyMatrix", {nS, R}, sizeof(uint32_t)}};
    for (const Check &c : checks) {
        if ((rc = dataset_info(h, c.name, d))) return rc;
        if (!dims_are(d, c.dims, c.size)) return fail(LMXIO_ERR_FORMAT, std::string(c.name) + ": invalid dataset dimensions");
    }
    if (info.hasRateNoise) {
        if ((rc = dataset_info(h, "/Model/Reaction/ReactionRateNoise", d))) return rc;
        if (!dims_are(d, {R, kNoiseCols}, sizeof(double)))
            return fail(LMXIO_ERR_FORMAT, "/Model/Reaction/ReactionRateNoise: invalid dataset dimensions");
    }
    bool ok = true;
    if (types) ok &= H5LTread_dataset(h, "/Model/Reaction/ReactionTypes", H5T_NATIVE_UINT32, types) >= 0;
    if (K) ok &= H5LTread_dataset_double(h, "/Model/Reaction/ReactionRateConstants", K) >= 0;
    if (noise) {
        if (info.hasRateNoise) ok &= H5LTread_dataset_double(h, "/Model/Reaction/ReactionRateNoise", noise) >= 0;
        else std::fill(noise, noise + (size_t)R * kNoiseCols, 0.0);
    }
    if (S) ok &= H5LTread_dataset(h, "/Model/Reaction/StoichiometricMatrix", H5T_NATIVE_INT32, S) >= 0;
    if (D) ok &= H5LTread_dataset(h, "/Model/Reaction/DependencyMatrix", H5T_NATIVE_UINT32, D) >= 0;
    return ok ? LMXIO_OK : fail(LMXIO_ERR_IO, "cannot read the reaction model");
}

int lmxio_set_reaction_model(lmxio_file *f, uint32_t nS, uint32_t R, const uint32_t *initialCounts,
                             const uint32_t *types, const double *K, const double *noise, const int32_t *S,
                             const uint32_t *D)
{
    if (!f || !initialCounts || nS == 0 || (R && (!types || !K || !S || !D)))
        return fail(LMXIO_ERR_ARG, "invalid reaction model arguments");
    if (!f->writable) return fail(LMXIO_ERR_ARG, "file opened read-only");
    const hid_t h = f->file;
    if (H5Lexists(h, "/Model/Reaction", H5P_DEFAULT) > 0 && H5Ldelete(h, "/Model/Reaction", H5P_DEFAULT) < 0)
        return fail(LMXIO_ERR_IO, "cannot replace /Model/Reaction");
    int rc = make_group(h, "/Model/Reaction");
    if (rc) return rc;
    const char *g = "/Model/Reaction";
    bool ok = H5LTset_attribute_uint(h, g, "numberSpecies", &nS, 1) >= 0 &&
              H5LTset_attribute_uint(h, g, "numberReactions", &R, 1) >= 0;
    hsize_t dims[2] = {nS, 0};
    ok = ok && H5LTmake_dataset(h, "/Model/Reaction/InitialSpeciesCounts", 1, dims, H5T_STD_U32LE, initialCounts) >= 0;
    if (ok && R) {
        // The noise table is written only when some reaction has noise
        // (SimulationFile.cpp:520-538), as (variance, tau) rows.
        bool any_noise = false;
        if (noise)
            for (size_t i = 0; i < (size_t)R * kNoiseCols; i++) any_noise |= noise[i] != 0.0;
        dims[0] = R;
        ok = H5LTmake_dataset(h, "/Model/Reaction/ReactionTypes", 1, dims, H5T_STD_U32LE, types) >= 0;
        dims[1] = kRateCols;
        ok = ok && H5LTmake_dataset(h, "/Model/Reaction/ReactionRateConstants", 2, dims, H5T_IEEE_F64LE, K) >= 0;
        if (any_noise) {
            dims[1] = kNoiseCols;
            ok = ok && H5LTmake_dataset(h, "/Model/Reaction/ReactionRateNoise", 2, dims, H5T_IEEE_F64LE, noise) >= 0;
        }
        dims[0] = nS;
        dims[1] = R;
        ok = ok && H5LTmake_dataset(h, "/Model/Reaction/StoichiometricMatrix", 2, dims, H5T_STD_I32LE, S) >= 0;
        ok = ok && H5LTmake_dataset(h, "/Model/Reaction/DependencyMatrix", 2, dims, H5T_STD_U32LE, D) >= 0;
    }
    return ok ? LMXIO_OK : fail(LMXIO_ERR_IO, "cannot write the reaction model");
}

int lmxio_get_diffusion_info(lmxio_file *f, lmxio_diffusion_info *info)
{
    if (!f || !info) return fail(LMXIO_ERR_ARG, "null argument");
    std::memset(info, 0, sizeof(*info));
    const hid_t h = f->file;
    const char *g = "/Model/Diffusion";
    if (!link_exists(h, g)) return LMXIO_OK;
    info->present = 1;
    // The counts are optional here (older files carry only the lattice
    // geometry); the matrices that need them fail to read without them.
    bool ok = true;
    if (attr_exists(h, g, "numberSpecies")) ok &= H5LTget_attribute_uint(h, g, "numberSpecies", &info->numberSpecies) >= 0;
    if (attr_exists(h, g, "numberReactions")) ok &= H5LTget_attribute_uint(h, g, "numberReactions", &info->numberReactions) >= 0;
    if (attr_exists(h, g, "numberSiteTypes")) ok &= H5LTget_attribute_uint(h, g, "numberSiteTypes", &info->numberSiteTypes) >= 0;
    ok &= H5LTget_attribute_double(h, g, "latticeSpacing", &info->latticeSpacing) >= 0;
    ok &= H5LTget_attribute_uint(h, g, "latticeXSize", &info->X) >= 0;
    ok &= H5LTget_attribute_uint(h, g, "latticeYSize", &info->Y) >= 0;
    ok &= H5LTget_attribute_uint(h, g, "latticeZSize", &info->Z) >= 0;
    ok &= H5LTget_attribute_uint(h, g, "particlesPerSite", &info->P) >= 0;
    if (!ok) return fail(LMXIO_ERR_FORMAT, "/Model/Diffusion: missing or unreadable lattice attributes");
    info->bytesPerParticle = 1;  // an absent attribute means a byte lattice (SimulationFile.cpp:577-581)
    if (attr_exists(h, g, "bytes_per_particle") && H5LTget_attribute_uint(h, g, "bytes_per_particle", &info->bytesPerParticle) < 0)
        return fail(LMXIO_ERR_FORMAT, "/Model/Diffusion: unreadable bytes_per_particle");
    return LMXIO_OK;
}

int lmxio_get_diffusion_model(lmxio_file *f, double *DF, uint32_t *RL, void *lattice, size_t latticeBytes,
                              uint8_t *sites, size_t sitesBytes)
{
    lmxio_diffusion_info info;
    int rc = lmxio_get_diffusion_info(f, &info);
    if (rc) return rc;
    if (!info.present) return fail(LMXIO_ERR_FORMAT, "the file has no diffusion model");
    const hid_t h = f->file;
    const uint32_t nS = info.numberSpecies, nST = info.numberSiteTypes, R = info.numberReactions;
    Ds d;
    if (DF) {
        if ((rc = dataset_info(h, "/Model/Diffusion/DiffusionMatrix", d))) return rc;
        if (!nS || !nST || !dims_are(d, {nST, nST, nS}, sizeof(double)))
            return fail(LMXIO_ERR_FORMAT, "/Model/Diffusion/DiffusionMatrix: invalid dataset dimensions");
        if (H5LTread_dataset_double(h, "/Model/Diffusion/DiffusionMatrix", DF) < 0)
            return fail(LMXIO_ERR_IO, "cannot read DiffusionMatrix");
    }
    if (RL && R) {
        if ((rc = dataset_info(h, "/Model/Diffusion/ReactionLocationMatrix", d))) return rc;
        if (!dims_are(d, {R, nST}, sizeof(uint32_t)))
            return fail(LMXIO_ERR_FORMAT, "/Model/Diffusion/ReactionLocationMatrix: invalid dataset dimensions");
        if (H5LTread_dataset(h, "/Model/Diffusion/ReactionLocationMatrix", H5T_NATIVE_UINT32, RL) < 0)
            return fail(LMXIO_ERR_IO, "cannot read ReactionLocationMatrix");
    }
    if (lattice) {
        hid_t ft, mt;
        if ((rc = lattice_types(info.bytesPerParticle, &ft, &mt))) return fail(LMXIO_ERR_FORMAT, "unknown bytes per particle");
        if ((rc = dataset_info(h, "/Model/Diffusion/Lattice", d))) return rc;
        // The stored lattice is (x, y, z, p) row-major (SimulationFile.cpp:619-645).
        if (!dims_are(d, {info.X, info.Y, info.Z, info.P}, info.bytesPerParticle))
            return fail(LMXIO_ERR_FORMAT, "/Model/Diffusion/Lattice: invalid dataset dimensions");
        if (d.elements() * info.bytesPerParticle != latticeBytes) return fail(LMXIO_ERR_ARG, "incorrect lattice size");
        if (H5LTread_dataset(h, "/Model/Diffusion/Lattice", mt, lattice) < 0) return fail(LMXIO_ERR_IO, "cannot read Lattice");
    }
    if (sites) {
        if ((rc = dataset_info(h, "/Model/Diffusion/LatticeSites", d))) return rc;
        if (!dims_are(d, {info.X, info.Y, info.Z}, 1))
            return fail(LMXIO_ERR_FORMAT, "/Model/Diffusion/LatticeSites: invalid dataset dimensions");
        if (d.elements() != sitesBytes) return fail(LMXIO_ERR_ARG, "incorrect lattice sites size");
        if (H5LTread_dataset(h, "/Model/Diffusion/LatticeSites", H5T_NATIVE_UINT8, sites) < 0)
            return fail(LMXIO_ERR_IO, "cannot read LatticeSites");
    }
    return LMXIO_OK;
}

int lmxio_set_diffusion_model(lmxio_file *f, uint32_t nS, uint32_t R, uint32_t nST, const double *DF,
                              const uint32_t *RL, double latticeSpacing, uint32_t X, uint32_t Y, uint32_t Z,
                              uint32_t P, uint32_t bytesPerParticle, const void *lattice, const uint8_t *sites)
{
    if (!f || !DF || (R && !RL) || !nS || !nST || !X || !Y || !Z || !P || !(latticeSpacing > 0))
        return fail(LMXIO_ERR_ARG, "invalid diffusion model arguments");
    hid_t ft, mt;
    int rc = lattice_types(bytesPerParticle, &ft, &mt);
    if (rc) return rc;
    if (!f->writable) return fail(LMXIO_ERR_ARG, "file opened read-only");
    const hid_t h = f->file;
    if (H5Lexists(h, "/Model/Diffusion", H5P_DEFAULT) > 0 && H5Ldelete(h, "/Model/Diffusion", H5P_DEFAULT) < 0)
        return fail(LMXIO_ERR_IO, "cannot replace /Model/Diffusion");
    if ((rc = make_group(h, "/Model/Diffusion"))) return rc;
    const char *g = "/Model/Diffusion";
    bool ok = H5LTset_attribute_uint(h, g, "numberSpecies", &nS, 1) >= 0 &&
              H5LTset_attribute_uint(h, g, "numberReactions", &R, 1) >= 0 &&
              H5LTset_attribute_uint(h, g, "numberSiteTypes", &nST, 1) >= 0 &&
              H5LTset_attribute_double(h, g, "latticeSpacing", &latticeSpacing, 1) >= 0 &&
              H5LTset_attribute_uint(h, g, "latticeXSize", &X, 1) >= 0 &&
              H5LTset_attribute_uint(h, g, "latticeYSize", &Y, 1) >= 0 &&
              H5LTset_attribute_uint(h, g, "latticeZSize", &Z, 1) >= 0 &&
              H5LTset_attribute_uint(h, g, "particlesPerSite", &P, 1) >= 0 &&
              H5LTset_attribute_uint(h, g, "bytes_per_particle", &bytesPerParticle, 1) >= 0;
    hsize_t dims[4] = {nST, nST, nS, 0};
    ok = ok && H5LTmake_dataset(h, "/Model/Diffusion/DiffusionMatrix", 3, dims, H5T_IEEE_F64LE, DF) >= 0;
    if (ok && R) {
        dims[0] = R;
        dims[1] = nST;
        ok = H5LTmake_dataset(h, "/Model/Diffusion/ReactionLocationMatrix", 2, dims, H5T_STD_U32LE, RL) >= 0;
    }
    if (!ok) return fail(LMXIO_ERR_IO, "cannot write the diffusion model");
    return lmxio_set_diffusion_model_lattice(f, X, Y, Z, P, bytesPerParticle, lattice, sites);
}

int lmxio_set_diffusion_model_lattice(lmxio_file *f, uint32_t X, uint32_t Y, uint32_t Z, uint32_t P,
                                      uint32_t bytesPerParticle, const void *lattice, const uint8_t *sites)
{
    if (!f || !X || !Y || !Z || !P) return fail(LMXIO_ERR_ARG, "invalid lattice arguments");
    hid_t ft, mt;
    int rc = lattice_types(bytesPerParticle, &ft, &mt);
    if (rc) return rc;
    if (!f->writable) return fail(LMXIO_ERR_ARG, "file opened read-only");
    const char *g = "/Model/Diffusion";
    if (!link_exists(f->file, g)) return fail(LMXIO_ERR_FORMAT, "the file has no diffusion model");
    const hid_t dg = H5Gopen2(f->file, g, H5P_DEFAULT);
    if (dg < 0) return fail(LMXIO_ERR_IO, "cannot open /Model/Diffusion");
    const hsize_t ldims[4] = {X, Y, Z, P};
    if (lattice && H5Lexists(dg, "Lattice", H5P_DEFAULT) > 0) H5Ldelete(dg, "Lattice", H5P_DEFAULT);
    if (sites && H5Lexists(dg, "LatticeSites", H5P_DEFAULT) > 0) H5Ldelete(dg, "LatticeSites", H5P_DEFAULT);
    if (lattice) rc = write_frame(dg, "Lattice", 4, ldims, ft, mt, lattice);
    if (!rc && sites) rc = write_frame(dg, "LatticeSites", 3, ldims, H5T_STD_U8LE, H5T_NATIVE_UINT8, sites);
    H5Gclose(dg);
    return rc;
}

int lmxio_set_empty_spatial_model(lmxio_file *f)
{
    if (!f) return fail(LMXIO_ERR_ARG, "null argument");
    if (!f->writable) return fail(LMXIO_ERR_ARG, "file opened read-only");
    const hid_t h = f->file;
    if (H5Lexists(h, "/Model/Spatial", H5P_DEFAULT) > 0 && H5Ldelete(h, "/Model/Spatial", H5P_DEFAULT) < 0)
        return fail(LMXIO_ERR_IO, "cannot replace /Model/Spatial");
    int rc = make_group(h, "/Model/Spatial");
    if (!rc) rc = make_group(h, "/Model/Spatial/Regions");
    if (!rc) rc = make_group(h, "/Model/Spatial/Obstacles");
    const unsigned int zero = 0;
    if (!rc && (H5LTset_attribute_uint(h, "/Model/Spatial/Regions", "numberRegions", &zero, 1) < 0 ||
                H5LTset_attribute_uint(h, "/Model/Spatial/Obstacles", "numberObstacles", &zero, 1) < 0))
        rc = fail(LMXIO_ERR_IO, "cannot write the spatial model");
    return rc;
}

int lmxio_append_species_counts(lmxio_file *f, uint32_t replicate, uint32_t n, uint32_t nS, const double *times,
                                const uint32_t *counts)
{
    if (!f || (n && (!times || !counts)) || !nS) return fail(LMXIO_ERR_ARG, "invalid species-count arguments");
    if (!f->writable) return fail(LMXIO_ERR_ARG, "file opened read-only");
    hid_t g;
    int rc = open_replicate(f, replicate, true, &g);
    if (rc) return rc;
    rc = append_rows(g, "SpeciesCounts", H5T_NATIVE_UINT32, nS, n, counts);
    if (!rc) rc = append_rows(g, "SpeciesCountTimes", H5T_NATIVE_DOUBLE, 0, n, times);
    H5Gclose(g);
    return rc;
}

int lmxio_append_lattice(lmxio_file *f, uint32_t replicate, double time, uint32_t X, uint32_t Y, uint32_t Z,
                         uint32_t P, uint32_t bytesPerParticle, const void *lattice)
{
    if (!f || !lattice || !X || !Y || !Z || !P) return fail(LMXIO_ERR_ARG, "invalid lattice arguments");
    hid_t ft, mt;
    int rc = lattice_types(bytesPerParticle, &ft, &mt);
    if (rc) return rc;
    const hsize_t dims[4] = {X, Y, Z, P};
    return append_frame(f, replicate, "Lattice", "LatticeTimes", time, 4, dims, ft, mt, lattice);
}

int lmxio_append_sites(lmxio_file *f, uint32_t replicate, double time, uint32_t X, uint32_t Y, uint32_t Z,
                       const uint8_t *sites)
{
    if (!f || !sites || !X || !Y || !Z) return fail(LMXIO_ERR_ARG, "invalid sites arguments");
    const hsize_t dims[3] = {X, Y, Z};
    return append_frame(f, replicate, "Sites", "SiteTimes", time, 3, dims, H5T_STD_U8LE, H5T_NATIVE_UINT8, sites);
}

int lmxio_replicate_exists(lmxio_file *f, uint32_t replicate, int *out)
{
    if (!f || !out) return fail(LMXIO_ERR_ARG, "null argument");
    *out = link_exists(f->file, replicate_name(replicate)) ? 1 : 0;
    return LMXIO_OK;
}

int lmxio_get_species_counts_size(lmxio_file *f, uint32_t replicate, uint32_t *n, uint32_t *nS)
{
    if (!f || !n || !nS) return fail(LMXIO_ERR_ARG, "null argument");
    const std::string rep = replicate_name(replicate);
    Ds d, t;
    int rc = dataset_info(f->file, rep + "/SpeciesCounts", d);
    if (!rc) rc = dataset_info(f->file, rep + "/SpeciesCountTimes", t);
    if (rc) return rc;
    if (d.dims.size() != 2 || d.size != sizeof(int32_t)) return fail(LMXIO_ERR_FORMAT, "SpeciesCounts: invalid dataset rank or type");
    if (t.dims.size() != 1 || t.size != sizeof(double) || t.dims[0] != d.dims[0])
        return fail(LMXIO_ERR_FORMAT, "SpeciesCountTimes: invalid dataset rank, type or length");
    *n = (uint32_t)d.dims[0];
    *nS = (uint32_t)d.dims[1];
    return LMXIO_OK;
}

int lmxio_get_species_counts(lmxio_file *f, uint32_t replicate, double *times, uint32_t *counts)
{
    uint32_t n = 0, nS = 0;
    int rc = lmxio_get_species_counts_size(f, replicate, &n, &nS);
    if (rc || n == 0) return rc;
    const std::string rep = replicate_name(replicate);
    if (counts && H5LTread_dataset(f->file, (rep + "/SpeciesCounts").c_str(), H5T_NATIVE_UINT32, counts) < 0)
        return fail(LMXIO_ERR_IO, "cannot read SpeciesCounts");
    if (times && H5LTread_dataset_double(f->file, (rep + "/SpeciesCountTimes").c_str(), times) < 0)
        return fail(LMXIO_ERR_IO, "cannot read SpeciesCountTimes");
    return LMXIO_OK;
}

int lmxio_get_lattice_count(lmxio_file *f, uint32_t replicate, uint32_t *n)
{
    if (!f || !n) return fail(LMXIO_ERR_ARG, "null argument");
    const std::string path = replicate_name(replicate) + "/LatticeTimes";
    *n = 0;
    if (!link_exists(f->file, path)) return LMXIO_OK;
    Ds d;
    int rc = dataset_info(f->file, path, d);
    if (rc) return rc;
    if (d.dims.size() != 1 || d.size != sizeof(double)) return fail(LMXIO_ERR_FORMAT, "LatticeTimes: invalid dataset rank or type");
    *n = (uint32_t)d.dims[0];
    return LMXIO_OK;
}

int lmxio_get_lattice(lmxio_file *f, uint32_t replicate, uint32_t index, double *time, void *lattice,
                      size_t latticeBytes, uint32_t *bytesPerParticle)
{
    uint32_t n = 0;
    int rc = lmxio_get_lattice_count(f, replicate, &n);
    if (rc) return rc;
    if (index >= n) return fail(LMXIO_ERR_ARG, "lattice index out of range");
    const std::string rep = replicate_name(replicate);
    char name[32];
    std::snprintf(name, sizeof(name), "/Lattice/%010u", index);
    const std::string path = rep + name;
    Ds d;
    if ((rc = dataset_info(f->file, path, d))) return rc;
    if (d.dims.size() != 4) return fail(LMXIO_ERR_FORMAT, path + ": invalid dataset rank");
    if (d.size != 1 && d.size != 4) return fail(LMXIO_ERR_FORMAT, path + ": unsupported particle size");
    if (bytesPerParticle) *bytesPerParticle = (uint32_t)d.size;
    if (lattice) {
        if (latticeBytes != d.elements() * d.size) return fail(LMXIO_ERR_ARG, "lattice buffer size differs from the stored frame");
        if (H5LTread_dataset(f->file, path.c_str(), d.size == 1 ? H5T_NATIVE_UINT8 : H5T_NATIVE_UINT32, lattice) < 0)
            return fail(LMXIO_ERR_IO, "cannot read " + path);
    }
    if (time) {
        std::vector<double> t(n);
        if (H5LTread_dataset_double(f->file, (rep + "/LatticeTimes").c_str(), t.data()) < 0)
            return fail(LMXIO_ERR_IO, "cannot read LatticeTimes");
        *time = t[index];
    }
    return LMXIO_OK;
}

int lmxio_new_replicate_dataset_u32(lmxio_file *f, uint32_t replicate, const char *name, uint32_t n,
                                    const uint32_t *data)
{
    if (!f || !name || !*name || std::strchr(name, '/') || !n || !data)
        return fail(LMXIO_ERR_ARG, "invalid dataset arguments");
    if (!f->writable) return fail(LMXIO_ERR_ARG, "file opened read-only");
    hid_t g;
    int rc = open_replicate(f, replicate, true, &g);
    if (rc) return rc;
    if (H5Lexists(g, name, H5P_DEFAULT) > 0) {
        rc = fail(LMXIO_ERR_ARG, std::string("dataset already exists: ") + name);
    } else {
        const hsize_t dims[1] = {n}, chunk[1] = {std::min<hsize_t>(kGzipChunk, n)};
        const hid_t sp = H5Screate_simple(1, dims, nullptr);
        const hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
        H5Pset_deflate(pl, kGzipLevel);
        H5Pset_chunk(pl, 1, chunk);
        const hid_t ds = H5Dcreate2(g, name, H5T_STD_U32LE, sp, H5P_DEFAULT, pl, H5P_DEFAULT);
        const herr_t st = ds < 0 ? -1 : H5Dwrite(ds, H5T_NATIVE_UINT32, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
        if (ds >= 0) H5Dclose(ds);
        H5Pclose(pl);
        H5Sclose(sp);
        if (st < 0) rc = fail(LMXIO_ERR_IO, std::string("cannot write ") + name);
    }
    H5Gclose(g);
    return rc;
}

int lmxio_get_replicate_dataset_u32(lmxio_file *f, uint32_t replicate, const char *name, uint32_t *data,
                                    uint32_t cap, uint32_t *n)
{
    if (!f || !name || !*name || !n) return fail(LMXIO_ERR_ARG, "null argument");
    const std::string path = replicate_name(replicate) + "/" + name;
    Ds d;
    int rc = dataset_info(f->file, path, d);
    if (rc) return rc;
    if (d.dims.size() != 1 || d.cls != H5T_INTEGER) return fail(LMXIO_ERR_FORMAT, path + ": not a rank-1 integer dataset");
    *n = (uint32_t)d.dims[0];
    if (data) {
        if (cap < *n) return fail(LMXIO_ERR_ARG, "buffer too small");
        if (H5LTread_dataset(f->file, path.c_str(), H5T_NATIVE_UINT32, data) < 0) return fail(LMXIO_ERR_IO, "cannot read " + path);
    }
    return LMXIO_OK;
}

}  // extern "C"
