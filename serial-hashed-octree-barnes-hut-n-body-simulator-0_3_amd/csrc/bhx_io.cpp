// bhx_io.cpp — snapshot files and visualisation export (host).
//
// Snapshots have no reference counterpart (the reference regenerates its
// bodies at every start).  The export writes, as offline legacy-VTK files,
// what the reference draws each frame: a sphere per body (VisualizeBodies,
// Body.hpp:326-336) and a box per node centred on the node centre with edge
// = node size (visualizeTree / visualizeNode, HashedOctree.cpp:939-974,
// ofDrawBox(center, size)).
//
// Both writers go through the public C ABI only (bhx_get_bodies,
// bhx_get_nodes, ...): no device code here.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <utility>
#include <vector>

#include "../../include/bhx.h"

namespace {

constexpr char kMagic[8] = {'B', 'H', 'X', 'S', 'N', 'A', 'P', '1'};
constexpr uint32_t kHeaderBytes = 256;
constexpr uint32_t kVersion = 1;

struct FileCloser {
    FILE* f = nullptr;
    ~FileCloser() {
        if (f) std::fclose(f);
    }
};

bool host_is_little_endian() {
    const uint16_t probe = 1;
    unsigned char b;
    std::memcpy(&b, &probe, 1);
    return b == 1;
}

void put_u32(unsigned char* b, uint32_t v) {
    for (int i = 0; i < 4; ++i) b[i] = (unsigned char)(v >> (8 * i));
}
void put_u64(unsigned char* b, uint64_t v) {
    for (int i = 0; i < 8; ++i) b[i] = (unsigned char)(v >> (8 * i));
}
void put_f64(unsigned char* b, double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    put_u64(b, u);
}
uint32_t get_u32(const unsigned char* b) {
    uint32_t v = 0;
    for (int i = 0; i < 4; ++i) v |= (uint32_t)b[i] << (8 * i);
    return v;
}
uint64_t get_u64(const unsigned char* b) {
    uint64_t v = 0;
    for (int i = 0; i < 8; ++i) v |= (uint64_t)b[i] << (8 * i);
    return v;
}
double get_f64(const unsigned char* b) {
    const uint64_t u = get_u64(b);
    double v;
    std::memcpy(&v, &u, 8);
    return v;
}

// count 8-byte words, stored little-endian
bool write_words(FILE* f, const void* data, size_t count) {
    if (host_is_little_endian()) return std::fwrite(data, 8, count, f) == count;
    const unsigned char* p = static_cast<const unsigned char*>(data);
    unsigned char w[8];
    for (size_t i = 0; i < count; ++i) {
        for (int k = 0; k < 8; ++k) w[k] = p[8 * i + 7 - k];
        if (std::fwrite(w, 8, 1, f) != 1) return false;
    }
    return true;
}
bool read_words(FILE* f, void* data, size_t count) {
    if (std::fread(data, 8, count, f) != count) return false;
    if (!host_is_little_endian()) {
        unsigned char* p = static_cast<unsigned char*>(data);
        for (size_t i = 0; i < count; ++i)
            for (int k = 0; k < 4; ++k) std::swap(p[8 * i + k], p[8 * i + 7 - k]);
    }
    return true;
}
bool skip_words(FILE* f, size_t count) { return std::fseek(f, (long)(8 * count), SEEK_CUR) == 0; }

void encode_header(unsigned char* h, const bhx_snapshot_info& in, uint64_t nsys, bool has_acc, bool has_comp) {
    std::memset(h, 0, kHeaderBytes);
    std::memcpy(h, kMagic, 8);
    put_u32(h + 8, kHeaderBytes);
    put_u32(h + 12, kVersion);
    put_u64(h + 16, in.n);
    put_u64(h + 24, nsys);
    put_f64(h + 32, in.time);
    put_u64(h + 40, in.step);
    put_u32(h + 48, (has_acc ? 1u : 0u) | (has_comp ? 2u : 0u));
    put_f64(h + 56, in.cfg.theta);
    put_f64(h + 64, in.cfg.eps);
    put_f64(h + 72, in.cfg.dt);
    const int32_t ints[10] = {in.cfg.leaf_capacity, in.cfg.mode,        in.cfg.precision, in.cfg.quadrupole,
                              in.cfg.permute_acc,   in.cfg.compensated, in.cfg.device,    in.cfg.group_walk,
                              in.cfg.geometry,      in.cfg.group_size};
    for (int k = 0; k < 10; ++k) put_u32(h + 80 + 4 * k, (uint32_t)ints[k]);
}

int decode_header(const unsigned char* h, long file_bytes, bhx_snapshot_info* out) {
    if (std::memcmp(h, kMagic, 8) != 0) return BHX_E_IO;
    if (get_u32(h + 8) != kHeaderBytes || get_u32(h + 12) != kVersion) return BHX_E_IO;
    bhx_snapshot_info in;
    std::memset(&in, 0, sizeof in);
    in.n = get_u64(h + 16);
    in.nsys = get_u64(h + 24);
    in.time = get_f64(h + 32);
    in.step = get_u64(h + 40);
    const uint32_t flags = get_u32(h + 48);
    if (flags & ~3u) return BHX_E_IO;
    in.has_acc = (int32_t)(flags & 1u);
    in.has_comp = (int32_t)((flags >> 1) & 1u);
    in.cfg.theta = get_f64(h + 56);
    in.cfg.eps = get_f64(h + 64);
    in.cfg.dt = get_f64(h + 72);
    int32_t* ints[10] = {&in.cfg.leaf_capacity, &in.cfg.mode,        &in.cfg.precision, &in.cfg.quadrupole,
                         &in.cfg.permute_acc,   &in.cfg.compensated, &in.cfg.device,    &in.cfg.group_walk,
                         &in.cfg.geometry,      &in.cfg.group_size};
    for (int k = 0; k < 10; ++k) *ints[k] = (int32_t)get_u32(h + 80 + 4 * k);   // older files: group_size 0
    // sizes must match the file exactly (guards against truncation and against
    // counts that would overflow the arithmetic below)
    if (in.n > ((uint64_t)1 << 40) || in.nsys == 0 || in.nsys > ((uint64_t)1 << 32)) return BHX_E_IO;
    const uint64_t words = 7 * in.n + (in.has_acc ? 3 * in.n : 0) + (in.has_comp ? 6 * in.n : 0) + in.nsys + 1;
    if (file_bytes < 0 || (uint64_t)file_bytes != (uint64_t)kHeaderBytes + 8 * words) return BHX_E_IO;
    *out = in;
    return BHX_OK;
}

int open_snapshot(const char* path, FileCloser& fc, bhx_snapshot_info* info) {
    fc.f = std::fopen(path, "rb");
    if (!fc.f) return BHX_E_IO;
    if (std::fseek(fc.f, 0, SEEK_END) != 0) return BHX_E_IO;
    const long bytes = std::ftell(fc.f);
    if (std::fseek(fc.f, 0, SEEK_SET) != 0) return BHX_E_IO;
    unsigned char h[kHeaderBytes];
    if (bytes < (long)kHeaderBytes || std::fread(h, 1, kHeaderBytes, fc.f) != kHeaderBytes) return BHX_E_IO;
    return decode_header(h, bytes, info);
}

// ---- legacy VTK: big-endian binary payloads
void be32(unsigned char* b, uint32_t v) {
    for (int i = 0; i < 4; ++i) b[i] = (unsigned char)(v >> (8 * (3 - i)));
}
bool write_be_floats(FILE* f, const std::vector<float>& v) {
    std::vector<unsigned char> buf(4 * v.size());
    for (size_t i = 0; i < v.size(); ++i) {
        uint32_t u;
        std::memcpy(&u, &v[i], 4);
        be32(&buf[4 * i], u);
    }
    return std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
}
bool write_be_ints(FILE* f, const std::vector<int32_t>& v) {
    std::vector<unsigned char> buf(4 * v.size());
    for (size_t i = 0; i < v.size(); ++i) be32(&buf[4 * i], (uint32_t)v[i]);
    return std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
}

}  // namespace

extern "C" {

int bhx_snapshot_write(const char* path, const bhx_snapshot_info* info, const double* pos, const double* vel,
                       const double* mass, const double* acc, const double* comp, const uint64_t* offsets) {
    if (!path || !info || !pos || !vel || !mass) return BHX_E_NULL;
    const uint64_t n = info->n;
    const uint64_t nsys = offsets ? info->nsys : 1;
    if (nsys == 0) return BHX_E_ARG;
    if (offsets) {
        if (offsets[0] != 0 || offsets[nsys] != n) return BHX_E_ARG;
        for (uint64_t k = 0; k < nsys; ++k)
            if (offsets[k + 1] < offsets[k]) return BHX_E_ARG;
    }
    FileCloser fc;
    fc.f = std::fopen(path, "wb");
    if (!fc.f) return BHX_E_IO;
    unsigned char h[kHeaderBytes];
    encode_header(h, *info, nsys, acc != nullptr, comp != nullptr);
    if (std::fwrite(h, 1, kHeaderBytes, fc.f) != kHeaderBytes) return BHX_E_IO;
    if (!write_words(fc.f, pos, 3 * n) || !write_words(fc.f, vel, 3 * n) || !write_words(fc.f, mass, n))
        return BHX_E_IO;
    if (acc && !write_words(fc.f, acc, 3 * n)) return BHX_E_IO;
    if (comp && !write_words(fc.f, comp, 6 * n)) return BHX_E_IO;
    if (offsets) {
        if (!write_words(fc.f, offsets, nsys + 1)) return BHX_E_IO;
    } else {
        const uint64_t one[2] = {0, n};
        if (!write_words(fc.f, one, 2)) return BHX_E_IO;
    }
    const int rc = std::fclose(fc.f);
    fc.f = nullptr;
    return rc == 0 ? BHX_OK : BHX_E_IO;
}

int bhx_snapshot_read_info(const char* path, bhx_snapshot_info* info) {
    if (!path || !info) return BHX_E_NULL;
    FileCloser fc;
    return open_snapshot(path, fc, info);
}

int bhx_snapshot_read(const char* path, double* pos, double* vel, double* mass, double* acc, double* comp,
                      uint64_t* offsets) {
    if (!path) return BHX_E_NULL;
    FileCloser fc;
    bhx_snapshot_info in;
    int rc = open_snapshot(path, fc, &in);
    if (rc != BHX_OK) return rc;
    const uint64_t n = in.n;
    auto part = [&](void* dst, size_t words) { return dst ? read_words(fc.f, dst, words) : skip_words(fc.f, words); };
    if (!part(pos, 3 * n) || !part(vel, 3 * n) || !part(mass, n)) return BHX_E_IO;
    if (in.has_acc && !part(acc, 3 * n)) return BHX_E_IO;
    if (in.has_comp && !part(comp, 6 * n)) return BHX_E_IO;
    if (!part(offsets, in.nsys + 1)) return BHX_E_IO;
    return BHX_OK;
}

int bhx_save(bhx_ctx* ctx, const char* path, double time, uint64_t step) {
    if (!ctx || !path) return BHX_E_NULL;
    bhx_snapshot_info in;
    std::memset(&in, 0, sizeof in);
    size_t n = 0, nsys = 0;
    int rc;
    if ((rc = bhx_num_bodies(ctx, &n)) != BHX_OK) return rc;
    if ((rc = bhx_num_systems(ctx, &nsys)) != BHX_OK) return rc;
    if ((rc = bhx_get_config(ctx, &in.cfg)) != BHX_OK) return rc;
    in.n = n;
    in.nsys = nsys ? nsys : 1;
    in.time = time;
    in.step = step;
    std::vector<double> pos(3 * n), vel(3 * n), mass(n), acc(3 * n), comp(6 * n);
    std::vector<uint64_t> off(in.nsys + 1);
    if (n && (rc = bhx_get_bodies(ctx, pos.data(), vel.data(), mass.data(), acc.data())) != BHX_OK) return rc;
    if (n && (rc = bhx_get_compensation(ctx, comp.data(), comp.data() + 3 * n)) != BHX_OK) return rc;
    if ((rc = bhx_get_systems(ctx, off.data())) != BHX_OK) return rc;
    return bhx_snapshot_write(path, &in, pos.data(), vel.data(), mass.data(), acc.data(), comp.data(), off.data());
}

int bhx_load(bhx_ctx* ctx, const char* path, double* time, uint64_t* step) {
    if (!ctx || !path) return BHX_E_NULL;
    bhx_snapshot_info in;
    int rc = bhx_snapshot_read_info(path, &in);
    if (rc != BHX_OK) return rc;
    std::vector<double> pos(3 * in.n), vel(3 * in.n), mass(in.n), acc(in.has_acc ? 3 * in.n : 0),
        comp(in.has_comp ? 6 * in.n : 0);
    std::vector<uint64_t> off(in.nsys + 1);
    rc = bhx_snapshot_read(path, pos.data(), vel.data(), mass.data(), in.has_acc ? acc.data() : nullptr,
                           in.has_comp ? comp.data() : nullptr, off.data());
    if (rc != BHX_OK) return rc;
    if ((rc = bhx_set_bodies(ctx, pos.data(), vel.data(), mass.data(), (size_t)in.n)) != BHX_OK) return rc;
    if (in.has_acc && in.n && (rc = bhx_set_accelerations(ctx, acc.data())) != BHX_OK) return rc;
    if (in.has_comp && in.n &&
        (rc = bhx_set_compensation(ctx, comp.data(), comp.data() + 3 * in.n)) != BHX_OK)
        return rc;
    if (in.nsys > 1 && (rc = bhx_set_systems(ctx, off.data(), (size_t)in.nsys)) != BHX_OK) return rc;
    if (time) *time = in.time;
    if (step) *step = in.step;
    return BHX_OK;
}

int bhx_vtk_write_bodies(const char* path, const double* pos, const double* vel, const double* mass, size_t n) {
    if (!path || (n && !pos)) return BHX_E_NULL;
    if (n > (size_t)INT32_MAX / 2) return BHX_E_ARG;   // the format's int32 vertex indices
    FileCloser fc;
    fc.f = std::fopen(path, "wb");
    if (!fc.f) return BHX_E_IO;
    FILE* f = fc.f;
    std::fprintf(f, "# vtk DataFile Version 3.0\nbhx bodies\nBINARY\nDATASET POLYDATA\nPOINTS %zu float\n", n);
    std::vector<float> pts(3 * n);
    for (size_t i = 0; i < 3 * n; ++i) pts[i] = (float)pos[i];
    if (!write_be_floats(f, pts)) return BHX_E_IO;
    std::fprintf(f, "\nVERTICES %zu %zu\n", n, 2 * n);
    std::vector<int32_t> verts(2 * n);
    for (size_t i = 0; i < n; ++i) {
        verts[2 * i] = 1;
        verts[2 * i + 1] = (int32_t)i;
    }
    if (!write_be_ints(f, verts)) return BHX_E_IO;
    std::fprintf(f, "\nPOINT_DATA %zu\nSCALARS mass float 1\nLOOKUP_TABLE default\n", n);
    std::vector<float> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = mass ? (float)mass[i] : 0.0f;
    if (!write_be_floats(f, v)) return BHX_E_IO;
    std::fprintf(f, "\nSCALARS speed float 1\nLOOKUP_TABLE default\n");
    for (size_t i = 0; i < n; ++i)
        v[i] = vel ? (float)std::sqrt(vel[3 * i] * vel[3 * i] + vel[3 * i + 1] * vel[3 * i + 1] +
                                      vel[3 * i + 2] * vel[3 * i + 2])
                   : 0.0f;
    if (!write_be_floats(f, v)) return BHX_E_IO;
    std::fprintf(f, "\n");
    const int rc = std::fclose(fc.f);
    fc.f = nullptr;
    return rc == 0 ? BHX_OK : BHX_E_IO;
}

int bhx_vtk_write_nodes(const char* path, const bhx_node* nodes, size_t count, int min_level, int max_level) {
    if (!path || (count && !nodes)) return BHX_E_NULL;
    if (min_level > max_level) return BHX_E_ARG;
    std::vector<const bhx_node*> sel;
    for (size_t i = 0; i < count; ++i)
        if (nodes[i].level >= min_level && nodes[i].level <= max_level) sel.push_back(&nodes[i]);
    const size_t m = sel.size();
    if (m > (size_t)INT32_MAX / 9) return BHX_E_ARG;   // the format's int32 point indices and cell list
    FileCloser fc;
    fc.f = std::fopen(path, "wb");
    if (!fc.f) return BHX_E_IO;
    FILE* f = fc.f;
    std::fprintf(f, "# vtk DataFile Version 3.0\nbhx nodes\nBINARY\nDATASET UNSTRUCTURED_GRID\nPOINTS %zu float\n",
                 8 * m);
    // VTK_VOXEL corner order: x fastest, then y, then z
    std::vector<float> pts(24 * m);
    for (size_t j = 0; j < m; ++j) {
        const bhx_node& nd = *sel[j];
        const double h = 0.5 * nd.size;
        for (int c = 0; c < 8; ++c) {
            pts[24 * j + 3 * c] = (float)(nd.center[0] + ((c & 1) ? h : -h));
            pts[24 * j + 3 * c + 1] = (float)(nd.center[1] + ((c & 2) ? h : -h));
            pts[24 * j + 3 * c + 2] = (float)(nd.center[2] + ((c & 4) ? h : -h));
        }
    }
    if (!write_be_floats(f, pts)) return BHX_E_IO;
    std::fprintf(f, "\nCELLS %zu %zu\n", m, 9 * m);
    std::vector<int32_t> cells(9 * m);
    for (size_t j = 0; j < m; ++j) {
        cells[9 * j] = 8;
        for (int c = 0; c < 8; ++c) cells[9 * j + 1 + c] = (int32_t)(8 * j + c);
    }
    if (!write_be_ints(f, cells)) return BHX_E_IO;
    std::fprintf(f, "\nCELL_TYPES %zu\n", m);
    if (!write_be_ints(f, std::vector<int32_t>(m, 11))) return BHX_E_IO;
    std::fprintf(f, "\nCELL_DATA %zu\nSCALARS level int 1\nLOOKUP_TABLE default\n", m);
    std::vector<int32_t> iv(m);
    for (size_t j = 0; j < m; ++j) iv[j] = sel[j]->level;
    if (!write_be_ints(f, iv)) return BHX_E_IO;
    std::fprintf(f, "\nSCALARS count int 1\nLOOKUP_TABLE default\n");
    for (size_t j = 0; j < m; ++j) iv[j] = sel[j]->n > INT32_MAX ? INT32_MAX : (int32_t)sel[j]->n;
    if (!write_be_ints(f, iv)) return BHX_E_IO;
    std::fprintf(f, "\nSCALARS mass float 1\nLOOKUP_TABLE default\n");
    std::vector<float> fv(m);
    for (size_t j = 0; j < m; ++j) fv[j] = (float)sel[j]->mass;
    if (!write_be_floats(f, fv)) return BHX_E_IO;
    std::fprintf(f, "\n");
    const int rc = std::fclose(fc.f);
    fc.f = nullptr;
    return rc == 0 ? BHX_OK : BHX_E_IO;
}

int bhx_export_vtk(bhx_ctx* ctx, const char* bodies_path, const char* nodes_path, int min_level, int max_level) {
    if (!ctx) return BHX_E_NULL;
    int rc;
    if (bodies_path) {
        size_t n = 0;
        if ((rc = bhx_num_bodies(ctx, &n)) != BHX_OK) return rc;
        std::vector<double> pos(3 * n), vel(3 * n), mass(n);
        if (n && (rc = bhx_get_bodies(ctx, pos.data(), vel.data(), mass.data(), nullptr)) != BHX_OK) return rc;
        if ((rc = bhx_vtk_write_bodies(bodies_path, pos.data(), vel.data(), mass.data(), n)) != BHX_OK) return rc;
    }
    if (nodes_path) {
        size_t m = 0;
        if ((rc = bhx_get_nodes(ctx, nullptr, 0, &m)) != BHX_OK) return rc;
        std::vector<bhx_node> nodes(m);
        if (m && (rc = bhx_get_nodes(ctx, nodes.data(), m, &m)) != BHX_OK) return rc;
        if ((rc = bhx_vtk_write_nodes(nodes_path, nodes.data(), m, min_level, max_level)) != BHX_OK) return rc;
    }
    return BHX_OK;
}

}  // extern "C"
