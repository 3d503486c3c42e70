// Output: the reference's NetCDF file (files.cpp:51-282, 592-700, 1205-1254, 1262-1299) written by this
// library's own classic-format (CDF-2, 64-bit offset) writer, and asynchronous staging of the output
// fields from HBM to pinned host memory, so a run writes output without a host round trip in its step
// loop (driver_serial.cpp:317-345).
//
// The file layout follows create_outfile (files.cpp:51-282) with basestate, modelBaseState and
// coordinates all true, as outfile_init creates it (files.cpp:1262-1285): dimensions x = NX, y = NY,
// z = NZ, t unlimited; record variables u-wind, v-wind, w-wind, theta, pi (t,x,y,z), with moisture
// rainfall (t,x,y) and qv, qc, qr (t,x,y,z); ubar ... pbar (x,y,z); tb, pib, qb, zu (z); lat (y),
// lon (x); time (t); topo (x,y); every variable NC_FLOAT.  The global attributes are create_outfile's,
// in its order, with last_file_successfully_written the restart status that file_output_status and
// get_file_output_status set and read (files.cpp:1205-1254).
//
// The classic format (Unidata's netCDF classic / 64-bit offset format specification): big-endian; a
// header of magic "CDF\2", numrecs, the dimension, global-attribute and variable lists (each variable
// with its vsize and the 8-byte offset `begin` of its data), every name and value padded to 4 bytes;
// then the non-record variables one after another, then the records, each holding one slice of every
// record variable in definition order.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "dyn_internal.hpp"

namespace dyn {
namespace {

enum : int32_t { kNcChar = 2, kNcInt = 4, kNcFloat = 5, kNcDouble = 6 };
enum : int32_t { kTagDimension = 0x0A, kTagVariable = 0x0B, kTagAttribute = 0x0C };

void put_be32(std::vector<uint8_t> &b, uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back(uint8_t(v >> s));
}
void put_be64(std::vector<uint8_t> &b, uint64_t v) {
    for (int s = 56; s >= 0; s -= 8) b.push_back(uint8_t(v >> s));
}
void put_name(std::vector<uint8_t> &b, const std::string &s) {
    put_be32(b, (uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
    while (b.size() % 4) b.push_back(0);
}
uint32_t get_be32(const uint8_t *p) { return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3]; }
uint64_t get_be64(const uint8_t *p) { return (uint64_t(get_be32(p)) << 32) | get_be32(p + 4); }
size_t type_size(int32_t t) { return t == kNcChar ? 1 : t == kNcInt || t == kNcFloat ? 4 : t == kNcDouble ? 8 : 0; }
size_t pad4(size_t n) { return (n + 3) & ~size_t(3); }

struct NcAtt {
    std::string name;
    int32_t type;
    uint32_t nelems;
    std::vector<uint8_t> values;  // big-endian, unpadded
};
struct NcDim {
    std::string name;
    uint32_t len;  // 0: the unlimited (record) dimension
};
struct NcVar {
    std::string name;
    std::vector<int32_t> dims;
    int32_t type = kNcFloat;
    uint64_t vsize = 0, begin = 0;
    uint64_t nvals = 0;  // values per record (record variables) or in all (the others)
    bool record = false;
};
struct NcFile {
    std::vector<NcDim> dims;
    std::vector<NcAtt> atts;
    std::vector<NcVar> vars;
    uint32_t numrecs = 0;
    uint64_t recsize = 0;
    size_t numrecs_offset = 4;
    size_t status_value_offset = 0;  // byte offset of last_file_successfully_written's value, 0 if absent
};

NcAtt att_float(const std::string &n, double v) {
    NcAtt a{n, kNcFloat, 1, {}};
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    put_be32(a.values, u);
    return a;
}
NcAtt att_int(const std::string &n, int v) {
    NcAtt a{n, kNcInt, 1, {}};
    put_be32(a.values, (uint32_t)v);
    return a;
}
NcAtt att_text(const std::string &n, const std::string &v) {
    NcAtt a{n, kNcChar, (uint32_t)v.size(), {}};
    a.values.assign(v.begin(), v.end());
    return a;
}

// vsize and begin of every variable: non-record data first, in definition order, then the records
std::vector<uint8_t> encode_header(NcFile &f) {
    auto build = [&](uint64_t data_start) {
        uint64_t off = data_start;
        for (auto &v : f.vars)
            if (!v.record) {
                v.begin = off;
                off += v.vsize;
            }
        f.recsize = 0;
        for (auto &v : f.vars)
            if (v.record) {
                v.begin = off + f.recsize;
                f.recsize += v.vsize;
            }
        std::vector<uint8_t> b = {'C', 'D', 'F', 2};
        put_be32(b, f.numrecs);
        put_be32(b, kTagDimension);
        put_be32(b, (uint32_t)f.dims.size());
        for (auto &d : f.dims) {
            put_name(b, d.name);
            put_be32(b, d.len);
        }
        put_be32(b, kTagAttribute);
        put_be32(b, (uint32_t)f.atts.size());
        for (auto &a : f.atts) {
            put_name(b, a.name);
            put_be32(b, (uint32_t)a.type);
            put_be32(b, a.nelems);
            if (a.name == "last_file_successfully_written") f.status_value_offset = b.size();
            b.insert(b.end(), a.values.begin(), a.values.end());
            while (b.size() % 4) b.push_back(0);
        }
        put_be32(b, kTagVariable);
        put_be32(b, (uint32_t)f.vars.size());
        for (auto &v : f.vars) {
            put_name(b, v.name);
            put_be32(b, (uint32_t)v.dims.size());
            for (int32_t d : v.dims) put_be32(b, (uint32_t)d);
            put_be32(b, 0);  // no variable attributes (ABSENT)
            put_be32(b, 0);
            put_be32(b, (uint32_t)v.type);
            put_be32(b, (uint32_t)std::min<uint64_t>(v.vsize, 0xffffffffu));
            put_be64(b, v.begin);
        }
        return b;
    };
    std::vector<uint8_t> h = build(0);  // the header's size does not depend on the offsets it holds
    return build(h.size());
}

void define_var(NcFile &f, const std::string &name, std::vector<int32_t> dims) {
    NcVar v;
    v.name = name;
    v.dims = dims;
    v.nvals = 1;
    for (int32_t d : dims) {
        if (f.dims[d].len == 0) v.record = true;
        else v.nvals *= f.dims[d].len;
    }
    v.vsize = pad4(v.nvals * type_size(v.type));
    f.vars.push_back(v);
}

// parse a classic (CDF-1 or CDF-2) header: enough of it to locate variables, numrecs and the status
bool parse_header(FILE *fp, NcFile &f, std::string &err) {
    std::vector<uint8_t> buf(1 << 16);
    size_t got = fread(buf.data(), 1, buf.size(), fp);
    size_t p = 0;
    auto need = [&](size_t n) {
        while (p + n > got) {
            buf.resize(buf.size() * 2);
            got += fread(buf.data() + got, 1, buf.size() - got, fp);
            if (p + n > got && feof(fp)) return false;
        }
        return true;
    };
    auto u32 = [&](uint32_t &v) {
        if (!need(4)) return false;
        v = get_be32(&buf[p]);
        p += 4;
        return true;
    };
    auto name = [&](std::string &s) {
        uint32_t n;
        if (!u32(n) || !need(pad4(n))) return false;
        s.assign((const char *)&buf[p], n);
        p += pad4(n);
        return true;
    };
    if (!need(8) || buf[0] != 'C' || buf[1] != 'D' || buf[2] != 'F' || (buf[3] != 1 && buf[3] != 2)) {
        err = "not a netCDF classic file";
        return false;
    }
    const bool cdf2 = buf[3] == 2;
    p = 4;
    uint32_t tag, n;
    if (!u32(f.numrecs)) return false;
    if (!u32(tag) || !u32(n)) return false;
    for (uint32_t d = 0; d < n; d++) {
        NcDim dim;
        if (!name(dim.name) || !u32(dim.len)) return false;
        f.dims.push_back(dim);
    }
    if (!u32(tag) || !u32(n)) return false;
    for (uint32_t a = 0; a < n; a++) {
        NcAtt att;
        uint32_t t;
        if (!name(att.name) || !u32(t) || !u32(att.nelems)) return false;
        att.type = (int32_t)t;
        size_t bytes = pad4(att.nelems * type_size(att.type));
        if (!need(bytes)) return false;
        if (att.name == "last_file_successfully_written" && att.type == kNcInt && att.nelems == 1) f.status_value_offset = p;
        att.values.assign(buf.begin() + p, buf.begin() + p + att.nelems * type_size(att.type));
        p += bytes;
        f.atts.push_back(att);
    }
    if (!u32(tag) || !u32(n)) return false;
    for (uint32_t v = 0; v < n; v++) {
        NcVar var;
        uint32_t nd, t;
        if (!name(var.name) || !u32(nd)) return false;
        var.nvals = 1;
        for (uint32_t d = 0; d < nd; d++) {
            uint32_t id;
            if (!u32(id) || id >= f.dims.size()) return false;
            var.dims.push_back((int32_t)id);
            if (f.dims[id].len == 0) var.record = true;
            else var.nvals *= f.dims[id].len;
        }
        uint32_t vtag, vn;
        if (!u32(vtag) || !u32(vn)) return false;
        for (uint32_t a = 0; a < vn; a++) {  // skip variable attributes
            std::string s;
            uint32_t at, ne;
            if (!name(s) || !u32(at) || !u32(ne)) return false;
            size_t bytes = pad4(ne * type_size((int32_t)at));
            if (!need(bytes)) return false;
            p += bytes;
        }
        uint32_t vs;
        if (!u32(t) || !u32(vs)) return false;
        var.type = (int32_t)t;
        var.vsize = vs;
        if (cdf2) {
            if (!need(8)) return false;
            var.begin = get_be64(&buf[p]);
            p += 8;
        } else {
            uint32_t b;
            if (!u32(b)) return false;
            var.begin = b;
        }
        f.vars.push_back(var);
    }
    f.recsize = 0;
    int nrec = 0;
    for (auto &v : f.vars)
        if (v.record) {
            f.recsize += v.vsize;
            nrec++;
        }
    if (nrec == 1)  // a lone record variable is not padded (the format's special case)
        for (auto &v : f.vars)
            if (v.record) f.recsize = v.nvals * type_size(v.type);
    return true;
}

int open_parsed(const char *path, const char *mode, FILE **fp, NcFile &f, const char *who) {
    *fp = fopen(path, mode);
    if (!*fp) {
        set_error(std::string(who) + ": cannot open " + path);
        return DYN_E_ARG;
    }
    std::string err = "truncated header";
    if (!parse_header(*fp, f, err)) {
        fclose(*fp);
        *fp = nullptr;
        set_error(std::string(who) + ": " + path + ": " + err);
        return DYN_E_ARG;
    }
    return DYN_OK;
}

const NcVar *find_var(const NcFile &f, const char *name) {
    for (auto &v : f.vars)
        if (v.name == name) return &v;
    return nullptr;
}

// values of one variable (one record of a record variable) written as NC_FLOAT from doubles
int write_values(FILE *fp, const NcFile &f, const NcVar &v, uint64_t record, const double *src, const float *srcf) {
    if (v.type != kNcFloat) {
        set_error("dyn_netcdf: only NC_FLOAT variables are written");
        return DYN_E_UNSUPPORTED;
    }
    const uint64_t off = v.begin + (v.record ? record * f.recsize : 0);
    if (fseeko(fp, (off_t)off, SEEK_SET) != 0) return DYN_E_ARG;
    std::vector<uint8_t> out(std::min<uint64_t>(v.nvals, 1 << 20) * 4);
    for (uint64_t i0 = 0; i0 < v.nvals; i0 += (1 << 20)) {
        const uint64_t n = std::min<uint64_t>(v.nvals - i0, 1 << 20);
        for (uint64_t i = 0; i < n; i++) {
            float x = src ? (float)src[i0 + i] : srcf[i0 + i];
            uint32_t u;
            memcpy(&u, &x, 4);
            out[4 * i] = uint8_t(u >> 24);
            out[4 * i + 1] = uint8_t(u >> 16);
            out[4 * i + 2] = uint8_t(u >> 8);
            out[4 * i + 3] = uint8_t(u);
        }
        if (fwrite(out.data(), 1, n * 4, fp) != n * 4) return DYN_E_ARG;
    }
    if (v.record && record + 1 > f.numrecs) {
        uint8_t be[4];
        const uint32_t nr = (uint32_t)(record + 1);
        for (int s = 0; s < 4; s++) be[s] = uint8_t(nr >> (24 - 8 * s));
        if (fseeko(fp, (off_t)f.numrecs_offset, SEEK_SET) != 0 || fwrite(be, 1, 4, fp) != 4) return DYN_E_ARG;
    }
    return DYN_OK;
}

}  // namespace

// ---- C ABI: the file ---------------------------------------------------------------------------------
extern "C" int dyn_netcdf_create(const char *path, int nx, int ny, int nz, int microphysics, const dyn_nc_meta *meta) {
    if (!path || nx < 1 || ny < 1 || nz < 1 || !meta) {
        set_error("dyn_netcdf_create: bad argument");
        return DYN_E_ARG;
    }
    NcFile f;
    f.dims = {{"x", (uint32_t)nx}, {"y", (uint32_t)ny}, {"z", (uint32_t)nz}, {"t", 0}};
    const int X = 0, Y = 1, Z = 2, TT = 3;
    // global attributes in create_outfile's order (files.cpp:84-116); a name set twice keeps one entry
    f.atts = {att_float("dx", meta->dx), att_float("dy", meta->dx), att_float("dz", meta->dz), att_float("dt", meta->dt),
              att_float("latoffset", meta->latoffset), att_float("lonoffset", meta->lonoffset),
              att_text("reanalysis_input_data", meta->reanalysis_input_data ? meta->reanalysis_input_data : ""),
              att_text("perturbation_input_data", meta->perturbation_input_data ? meta->perturbation_input_data : ""),
              att_text("basic_state_input_data", meta->basic_state_input_data ? meta->basic_state_input_data : ""),
              att_int("hydrostatic", 0), att_int("islinear", 0), att_int("use_linear_friction", 0),
              att_int("use_turbulent_stress", meta->use_turbulent_stress), att_int("microphysics_option", microphysics),
              att_int("use_terrain", 0), att_int("surface_heat_flux", 0), att_float("water_temperature", meta->water_temperature),
              att_int("use_landsea_from_file", 0), att_int("use_explicit_diffusion", 0), att_int("periodic_boundaries", 1),
              att_int("stretched_grid", 1), att_float("height_lowest_level", meta->height_lowest_level),
              att_int("index_lowest_level", 1), att_int("basic_state_option", meta->basic_state_option),
              att_int("perturbation_option", meta->perturbation_option), att_int("reanalysis_initialize_option", 1),
              att_int("last_file_successfully_written", DYN_NC_ALL_FILES_WRITTEN)};
    // variables in create_outfile's order (files.cpp:139-270)
    for (const char *n : {"u-wind", "v-wind", "w-wind", "theta", "pi"}) define_var(f, n, {TT, X, Y, Z});
    if (microphysics) {
        define_var(f, "rainfall", {TT, X, Y});
        for (const char *n : {"qv", "qc", "qr"}) define_var(f, n, {TT, X, Y, Z});
    }
    for (const char *n : {"ubar", "vbar", "wbar", "thbar", "qbar", "pbar"}) define_var(f, n, {X, Y, Z});
    for (const char *n : {"tb", "pib", "qb", "zu"}) define_var(f, n, {Z});
    define_var(f, "lat", {Y});
    define_var(f, "lon", {X});
    define_var(f, "time", {TT});
    define_var(f, "topo", {X, Y});
    std::vector<uint8_t> h = encode_header(f);
    FILE *fp = fopen(path, "wb");
    if (!fp) {
        set_error(std::string("dyn_netcdf_create: cannot create ") + path);
        return DYN_E_ARG;
    }
    bool ok = fwrite(h.data(), 1, h.size(), fp) == h.size();
    // non-record data, zero-filled until written (the topography of the flat path stays zero)
    uint64_t end = h.size();
    for (auto &v : f.vars)
        if (!v.record) end = std::max(end, v.begin + v.vsize);
    std::vector<uint8_t> zeros(1 << 20, 0);
    for (uint64_t off = h.size(); ok && off < end; off += zeros.size())
        ok = fwrite(zeros.data(), 1, (size_t)std::min<uint64_t>(zeros.size(), end - off), fp) == std::min<uint64_t>(zeros.size(), end - off);
    ok = (fclose(fp) == 0) && ok;
    if (!ok) {
        set_error(std::string("dyn_netcdf_create: write failed for ") + path);
        return DYN_E_ARG;
    }
    return DYN_OK;
}

extern "C" int dyn_netcdf_put(const char *path, const char *var, int record, const double *values, size_t count) {
    if (!path || !var || !values || record < 0) {
        set_error("dyn_netcdf_put: bad argument");
        return DYN_E_ARG;
    }
    NcFile f;
    FILE *fp = nullptr;
    if (int rc = open_parsed(path, "r+b", &fp, f, "dyn_netcdf_put")) return rc;
    const NcVar *v = find_var(f, var);
    int rc = DYN_OK;
    if (!v) {
        set_error(std::string("dyn_netcdf_put: no variable ") + var);
        rc = DYN_E_ARG;
    } else if (count != v->nvals) {
        set_error("dyn_netcdf_put: " + std::string(var) + " holds " + std::to_string(v->nvals) + " values per record");
        rc = DYN_E_ARG;
    } else {
        rc = write_values(fp, f, *v, (uint64_t)record, values, nullptr);
        if (rc && dyn_last_error()[0] == 0) set_error("dyn_netcdf_put: write failed");
    }
    if (fclose(fp) != 0 && !rc) rc = DYN_E_ARG;
    return rc;
}

extern "C" int dyn_netcdf_read(const char *path, const char *var, int record, double *values, size_t count) {
    if (!path || !var || !values || record < 0) {
        set_error("dyn_netcdf_read: bad argument");
        return DYN_E_ARG;
    }
    NcFile f;
    FILE *fp = nullptr;
    if (int rc = open_parsed(path, "rb", &fp, f, "dyn_netcdf_read")) return rc;
    const NcVar *v = find_var(f, var);
    int rc = DYN_OK;
    if (!v || count != v->nvals || (v->record && (uint32_t)record >= f.numrecs) ||
        (v->type != kNcFloat && v->type != kNcDouble && v->type != kNcInt)) {
        set_error(std::string("dyn_netcdf_read: ") + var + ": no such variable or record, or a count other than " +
                  std::to_string(v ? v->nvals : 0));
        rc = DYN_E_ARG;
    } else {
        const size_t ts = type_size(v->type);
        std::vector<uint8_t> raw(v->nvals * ts);
        const uint64_t off = v->begin + (v->record ? (uint64_t)record * f.recsize : 0);
        if (fseeko(fp, (off_t)off, SEEK_SET) != 0 || fread(raw.data(), 1, raw.size(), fp) != raw.size()) {
            set_error("dyn_netcdf_read: short file");
            rc = DYN_E_ARG;
        } else {
            for (uint64_t i = 0; i < v->nvals; i++) {
                const uint8_t *q = &raw[i * ts];
                if (v->type == kNcFloat) {
                    uint32_t u = get_be32(q);
                    float x;
                    memcpy(&x, &u, 4);
                    values[i] = x;
                } else if (v->type == kNcDouble) {
                    uint64_t u = get_be64(q);
                    double x;
                    memcpy(&x, &u, 8);
                    values[i] = x;
                } else {
                    values[i] = (double)(int32_t)get_be32(q);
                }
            }
        }
    }
    fclose(fp);
    return rc;
}

// file_output_status / get_file_output_status, files.cpp:1205-1254
extern "C" int dyn_netcdf_set_status(const char *path, int status) {
    NcFile f;
    FILE *fp = nullptr;
    if (int rc = open_parsed(path, "r+b", &fp, f, "dyn_netcdf_set_status")) return rc;
    int rc = DYN_OK;
    if (!f.status_value_offset) {
        set_error("dyn_netcdf_set_status: the file has no last_file_successfully_written attribute");
        rc = DYN_E_UNSUPPORTED;
    } else {
        uint8_t be[4] = {uint8_t((uint32_t)status >> 24), uint8_t((uint32_t)status >> 16), uint8_t((uint32_t)status >> 8),
                         uint8_t(status)};
        if (fseeko(fp, (off_t)f.status_value_offset, SEEK_SET) != 0 || fwrite(be, 1, 4, fp) != 4) rc = DYN_E_ARG;
    }
    if (fclose(fp) != 0 && !rc) rc = DYN_E_ARG;
    return rc;
}

extern "C" int dyn_netcdf_get_status(const char *path, int *status) {
    if (!status) return DYN_E_ARG;
    NcFile f;
    FILE *fp = nullptr;
    if (int rc = open_parsed(path, "rb", &fp, f, "dyn_netcdf_get_status")) return rc;
    *status = -1;  // absent (get_file_output_status returns -1)
    for (auto &a : f.atts)
        if (a.name == "last_file_successfully_written" && a.type == kNcInt && a.nelems == 1) *status = (int)get_be32(a.values.data());
    fclose(fp);
    return DYN_OK;
}

extern "C" int dyn_netcdf_records(const char *path) {
    NcFile f;
    FILE *fp = nullptr;
    if (int rc = open_parsed(path, "rb", &fp, f, "dyn_netcdf_records")) return rc;
    fclose(fp);
    return (int)f.numrecs;
}

// ---- staging of the output fields -----------------------------------------------------------------------
namespace {

// the fields output_meteorological_fields_to_file writes (files.cpp:683-698) and their file names
struct OutField {
    int field;
    const char *name;
};
constexpr OutField kOutDry[] = {{DYN_U, "u-wind"}, {DYN_V, "v-wind"}, {DYN_W, "w-wind"}, {DYN_PI, "pi"}, {DYN_TH, "theta"}};
constexpr OutField kOutMoist[] = {{DYN_QV, "qv"}, {DYN_QC, "qc"}, {DYN_QR, "qr"}, {DYN_ACCRAIN, "rainfall"}};

template <typename T>
__global__ void k_stage(float *__restrict__ dst, const T *__restrict__ src, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}

int out_fields(const Ctx *c, OutField *list) {
    int n = 0;
    for (auto &o : kOutDry) list[n++] = o;
    if (c->g.microphysics)
        for (auto &o : kOutMoist) list[n++] = o;
    return n;
}

void *field_device_ptr(Ctx *c, int field) {
    if (field >= DYN_U && field <= DYN_QR) return c->prog[field][DYN_ROLE_CUR];
    if (field == DYN_PI) return c->pi;
    if (field == DYN_ACCRAIN) return c->accrain;
    return nullptr;
}

size_t field_values(const Ctx *c, int field) { return field == DYN_ACCRAIN ? (size_t)c->g.nx * c->g.ny : c->npts; }

}  // namespace

int output_release(Ctx *c) {
    if (c->out.done) (void)hipEventSynchronize(c->out.done);
    if (c->out.dev) (void)hipFree(c->out.dev);
    if (c->out.host) (void)hipHostFree(c->out.host);
    if (c->out.done) (void)hipEventDestroy(c->out.done);
    if (c->out.snap) (void)hipEventDestroy(c->out.snap);
    if (c->out.copy) (void)hipStreamDestroy(c->out.copy);
    c->out = OutputStage{};
    return DYN_OK;
}

extern "C" int dyn_output_stage(dyn_ctx *ctx) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) {
        set_error("dyn_output_stage: null context");
        return DYN_E_ARG;
    }
    OutField list[16];
    const int n = out_fields(c, list);
    size_t total = 0;
    for (int i = 0; i < n; i++) total += field_values(c, list[i].field);
    OutputStage &o = c->out;
    if (!o.dev) {
        if (hipMalloc(&o.dev, total * sizeof(float)) != hipSuccess || hipHostMalloc(&o.host, total * sizeof(float)) != hipSuccess ||
            hipStreamCreateWithFlags(&o.copy, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&o.snap, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&o.done, hipEventDisableTiming) != hipSuccess) {
            output_release(c);
            set_error("dyn_output_stage: staging buffers could not be allocated");
            return DYN_E_HIP;
        }
        o.values = total;
    }
    // the previous copy to the host must have left the device buffer before the snapshot overwrites it
    if (o.pending) DYN_HIP(hipStreamWaitEvent(c->stream, o.done, 0));
    size_t off = 0;
    for (int i = 0; i < n; i++) {
        const size_t m = field_values(c, list[i].field);
        const void *src = field_device_ptr(c, list[i].field);
        const int blocks = (int)std::min<size_t>((m + 255) / 256, 4096);
        if (c->prec == DYN_FP32)
            hipLaunchKernelGGL(k_stage<float>, dim3(blocks), dim3(256), 0, c->stream, o.dev + off, (const float *)src, (long long)m);
        else
            hipLaunchKernelGGL(k_stage<double>, dim3(blocks), dim3(256), 0, c->stream, o.dev + off, (const double *)src, (long long)m);
        DYN_HIP(hipGetLastError());
        off += m;
    }
    // the copy runs on its own stream: the steps queued after this call do not wait for it
    DYN_HIP(hipEventRecord(o.snap, c->stream));
    DYN_HIP(hipStreamWaitEvent(o.copy, o.snap, 0));
    DYN_HIP(hipMemcpyAsync(o.host, o.dev, o.values * sizeof(float), hipMemcpyDeviceToHost, o.copy));
    DYN_HIP(hipEventRecord(o.done, o.copy));
    o.pending = true;
    return DYN_OK;
}

extern "C" int dyn_output_ready(dyn_ctx *ctx) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !c->out.pending) {
        set_error("dyn_output_ready: nothing staged");
        return DYN_E_STATE;
    }
    hipError_t e = hipEventQuery(c->out.done);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    set_error("dyn_output_ready: hipEventQuery failed");
    return DYN_E_HIP;
}

static int wait_staged(Ctx *c, const char *who) {
    if (!c || !c->out.pending) {
        set_error(std::string(who) + ": nothing staged");
        return DYN_E_STATE;
    }
    DYN_HIP(hipEventSynchronize(c->out.done));
    return DYN_OK;
}

extern "C" int dyn_output_get(dyn_ctx *ctx, int field, float *values, size_t count) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = wait_staged(c, "dyn_output_get")) return rc;
    OutField list[16];
    const int n = out_fields(c, list);
    size_t off = 0;
    for (int i = 0; i < n; i++) {
        const size_t m = field_values(c, list[i].field);
        if (list[i].field == field) {
            if (!values || count != m) {
                set_error("dyn_output_get: expected " + std::to_string(m) + " values");
                return DYN_E_ARG;
            }
            memcpy(values, c->out.host + off, m * sizeof(float));
            return DYN_OK;
        }
        off += m;
    }
    set_error("dyn_output_get: the field is not one of the output fields");
    return DYN_E_ARG;
}

extern "C" int dyn_netcdf_write_staged(dyn_ctx *ctx, const char *path, int record, double hours) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = wait_staged(c, "dyn_netcdf_write_staged")) return rc;
    NcFile f;
    FILE *fp = nullptr;
    if (int rc = open_parsed(path, "r+b", &fp, f, "dyn_netcdf_write_staged")) return rc;
    OutField list[16];
    const int n = out_fields(c, list);
    size_t off = 0;
    int rc = DYN_OK;
    for (int i = 0; i < n && !rc; i++) {
        const NcVar *v = find_var(f, list[i].name);
        const size_t m = field_values(c, list[i].field);
        if (!v || v->nvals != m) {
            set_error(std::string("dyn_netcdf_write_staged: the file has no matching ") + list[i].name);
            rc = DYN_E_ARG;
            break;
        }
        rc = write_values(fp, f, *v, (uint64_t)record, nullptr, c->out.host + off);
        off += m;
    }
    if (!rc) {  // write_time_to_file, files.cpp:628-660: hours = mtime / 3600
        const NcVar *t = find_var(f, "time");
        if (t) rc = write_values(fp, f, *t, (uint64_t)record, &hours, nullptr);
    }
    if (fclose(fp) != 0 && !rc) rc = DYN_E_ARG;
    if (rc && !dyn_last_error()[0]) set_error("dyn_netcdf_write_staged: write failed");
    return rc;
}

// outfile_init, files.cpp:1262-1285: create the file and write the time-independent data
extern "C" int dyn_netcdf_init(dyn_ctx *ctx, const char *path, const dyn_nc_meta *meta, const double *lat, const double *lon) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !path || !meta) {
        set_error("dyn_netcdf_init: bad argument");
        return DYN_E_ARG;
    }
    const dyn_grid &g = c->g;
    if (int rc = dyn_netcdf_create(path, g.nx, g.ny, g.nz, g.microphysics, meta)) return rc;
    std::vector<double> buf(c->npts);
    const int bars[] = {DYN_UBAR, DYN_VBAR, DYN_WBAR, DYN_THBAR, DYN_QBAR, DYN_PBAR};
    const char *names[] = {"ubar", "vbar", "wbar", "thbar", "qbar", "pbar"};
    for (int b = 0; b < 6; b++) {  // write_all_bvars, files.cpp:792-816
        if (bars[b] == DYN_QBAR && !g.microphysics) std::fill(buf.begin(), buf.end(), 0.0);
        else if (int rc = dyn_get_field(ctx, bars[b], 0, buf.data(), buf.size())) return rc;
        if (int rc = dyn_netcdf_put(path, names[b], 0, buf.data(), buf.size())) return rc;
    }
    // write_all_mbvars, files.cpp:824-851 (STRETCHED_GRID 1: zu holds zsu)
    const int cols[] = {DYN_COL_TB, DYN_COL_PIB, DYN_COL_QB, DYN_COL_ZSU};
    const char *cnames[] = {"tb", "pib", "qb", "zu"};
    for (int k = 0; k < 4; k++)
        if (int rc = dyn_netcdf_put(path, cnames[k], 0, c->colhost + (size_t)cols[k] * g.nz, (size_t)g.nz)) return rc;
    // write_all_coordinates, files.cpp:859-889
    std::vector<double> zy(g.ny, 0.0), zx(g.nx, 0.0);
    if (int rc = dyn_netcdf_put(path, "lat", 0, lat ? lat : zy.data(), (size_t)g.ny)) return rc;
    if (int rc = dyn_netcdf_put(path, "lon", 0, lon ? lon : zx.data(), (size_t)g.nx)) return rc;
    return DYN_OK;
}

}  // namespace dyn
