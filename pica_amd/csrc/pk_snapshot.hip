// pk_snapshot.hip -- raw snapshot files of a grid and of an ensemble: restart, fixtures and
// training data (no counterpart in pica, which has no I/O on this path).
//
// A file is one fixed 160-byte little-endian header (PkSnapHeader) followed by raw arrays:
//   grid:     PK_NCOMP components, each the whole storage (ghosts included) of dims[0] x dims[1]
//             x dims[2] elements, z fastest, in the grid's precision;
//   ensemble: nsct int32 per-supercell counts, then nfield arrays of `count` elements in
//             supercell order (the order the particle step sees):
//               PK_F64 (absolute positions): x, y, z, ux, uy, uz, w as float64;
//               PK_F32 (split positions):    cell (int32, linear physical cell), ox, oy, oz, ux,
//                                            uy, uz, w as float32.
// Loading restores the exact bits and order, so a run continued from a snapshot is identical to
// the uninterrupted run.  The target object must have the same geometry (checked).
#include <cstdio>
#include <cstring>
#include <vector>

#include "pk_internal.h"

namespace {

struct PkSnapHeader {
    char magic[8];          // "PKSNAP1" + NUL
    uint32_t kind;          // 1: grid, 2: ensemble
    uint32_t version;       // 1
    uint32_t prec;          // PK_F64 / PK_F32
    uint32_t posmode;       // ensemble: PK_POS_ABS / PK_POS_SPLIT
    int32_t dims[3];        // grid: storage size incl. ghosts; ensemble: cells
    int32_t ghost;          // grid: ghost width; ensemble: species index
    int32_t ncps[3];        // ensemble: cells per supercell
    int32_t nfield;         // grid: PK_NCOMP; ensemble: 7 or 8
    double origin[3];       // grid
    double step[3];
    int64_t count;          // grid: elements per component; ensemble: particles
    int64_t nsct;           // ensemble: supercells
    uint32_t esize;         // bytes per element
    uint32_t pad;
    int64_t resorts;        // ensemble: re-sorts done (the interleave period counts them)
    uint32_t reserved[6];
};
static_assert(sizeof(PkSnapHeader) == 160, "snapshot header is 160 bytes");

constexpr uint32_t SNAP_GRID = 1, SNAP_ENSEMBLE = 2, SNAP_VERSION = 1;
const char SNAP_MAGIC[8] = {'P', 'K', 'S', 'N', 'A', 'P', '1', 0};

struct File {
    FILE* f = nullptr;
    File(const char* path, const char* mode) : f(path ? std::fopen(path, mode) : nullptr) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
    bool write(const void* p, size_t bytes) { return bytes == 0 || std::fwrite(p, 1, bytes, f) == bytes; }
    bool read(void* p, size_t bytes) { return bytes == 0 || std::fread(p, 1, bytes, f) == bytes; }
};

PkSnapHeader header(uint32_t kind)
{
    PkSnapHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, SNAP_MAGIC, 8);
    h.kind = kind;
    h.version = SNAP_VERSION;
    return h;
}

int read_header(pk_ctx* c, File& in, uint32_t kind, PkSnapHeader& h)
{
    if (!in.f) return pk_fail(c, PK_E_ARG, "snapshot: cannot open the file");
    if (!in.read(&h, sizeof(h)) || std::memcmp(h.magic, SNAP_MAGIC, 8) != 0)
        return pk_fail(c, PK_E_ARG, "snapshot: not a snapshot file");
    if (h.version != SNAP_VERSION) return pk_fail(c, PK_E_UNSUPPORTED, "snapshot: unknown version");
    if (h.kind != kind) return pk_fail(c, PK_E_ARG, "snapshot: the file holds another kind of object");
    return PK_OK;
}

}  // namespace

extern "C" {

int pk_grid_save(pk_grid* g, const char* path)
{
    if (!g || !path) return PK_E_ARG;
    pk_ctx* c = g->ctx;
    PkSnapHeader h = header(SNAP_GRID);
    h.prec = (uint32_t)g->prec;
    for (int d = 0; d < 3; d++) {
        h.dims[d] = g->s[d];
        h.origin[d] = g->origin[d];
        h.step[d] = g->step[d];
    }
    h.ghost = g->ghost;
    h.nfield = PK_NCOMP;
    h.count = (int64_t)g->n;
    h.esize = g->prec == PK_F64 ? 8 : 4;
    std::vector<char> buf(g->n * h.esize);
    File out(path, "wb");
    if (!out.f) return pk_fail(c, PK_E_ARG, "pk_grid_save: cannot create the file");
    if (!out.write(&h, sizeof(h))) return pk_fail(c, PK_E_STATE, "pk_grid_save: write failed");
    PK_HIP(c, hipStreamSynchronize(c->stream));
    for (int comp = 0; comp < PK_NCOMP; comp++) {
        const void* src = g->comp(comp);
        PK_HIP(c, hipMemcpy(buf.data(), src, buf.size(), hipMemcpyDeviceToHost));
        if (!out.write(buf.data(), buf.size())) return pk_fail(c, PK_E_STATE, "pk_grid_save: write failed");
    }
    return PK_OK;
}

int pk_grid_load(pk_grid* g, const char* path)
{
    if (!g || !path) return PK_E_ARG;
    pk_ctx* c = g->ctx;
    File in(path, "rb");
    PkSnapHeader h;
    if (int r = read_header(c, in, SNAP_GRID, h)) return r;
    bool same = h.prec == (uint32_t)g->prec && h.ghost == g->ghost && h.nfield == PK_NCOMP &&
                h.count == (int64_t)g->n && h.esize == (g->prec == PK_F64 ? 8u : 4u);
    for (int d = 0; d < 3; d++)
        same = same && h.dims[d] == g->s[d] && h.origin[d] == g->origin[d] && h.step[d] == g->step[d];
    if (!same) return pk_fail(c, PK_E_ARG, "pk_grid_load: the snapshot's grid geometry or precision differs");
    std::vector<char> buf(g->n * h.esize);
    PK_HIP(c, hipStreamSynchronize(c->stream));
    for (int comp = 0; comp < PK_NCOMP; comp++) {
        if (!in.read(buf.data(), buf.size())) return pk_fail(c, PK_E_ARG, "pk_grid_load: truncated file");
        void* dst = g->comp(comp);
        PK_HIP(c, hipMemcpy(dst, buf.data(), buf.size(), hipMemcpyHostToDevice));
    }
    return PK_OK;
}

int pk_ens_save(pk_ens* e, const char* path)
{
    if (!e || !path) return PK_E_ARG;
    pk_ctx* c = e->ctx;
    if (int r = pk_ens_resort(e)) return r;
    const long nsct = e->nsct;
    std::vector<int32_t> cnt(nsct);
    std::vector<int64_t> off(nsct + 1);
    PK_HIP(c, hipStreamSynchronize(c->stream));
    PK_HIP(c, hipMemcpy(cnt.data(), e->cnt, sizeof(int32_t) * nsct, hipMemcpyDeviceToHost));
    PK_HIP(c, hipMemcpy(off.data(), e->off, sizeof(int64_t) * (nsct + 1), hipMemcpyDeviceToHost));
    int64_t np = 0;
    for (long s = 0; s < nsct; s++) np += cnt[s];
    PkSnapHeader h = header(SNAP_ENSEMBLE);
    h.prec = (uint32_t)e->prec;
    h.posmode = (uint32_t)e->posmode;
    for (int d = 0; d < 3; d++) {
        h.dims[d] = e->ncell[d];
        h.ncps[d] = e->ncps[d];
        h.step[d] = e->step[d];
    }
    h.ghost = e->type;
    h.nfield = e->nfield();
    h.count = np;
    h.nsct = nsct;
    h.esize = (uint32_t)e->esize();
    h.resorts = e->resorts;
    File out(path, "wb");
    if (!out.f) return pk_fail(c, PK_E_ARG, "pk_ens_save: cannot create the file");
    if (!out.write(&h, sizeof(h)) || !out.write(cnt.data(), sizeof(int32_t) * nsct))
        return pk_fail(c, PK_E_STATE, "pk_ens_save: write failed");
    // one field at a time: the capacity layout from the device, packed per supercell
    std::vector<char> full(e->capacity * h.esize), packed((size_t)np * h.esize);
    for (int f = 0; f < h.nfield; f++) {
        PK_HIP(c, hipMemcpy(full.data(), e->field(e->main, f), full.size(), hipMemcpyDeviceToHost));
        size_t o = 0;
        for (long s = 0; s < nsct; s++) {
            std::memcpy(packed.data() + o, full.data() + (size_t)off[s] * h.esize, (size_t)cnt[s] * h.esize);
            o += (size_t)cnt[s] * h.esize;
        }
        if (!out.write(packed.data(), packed.size())) return pk_fail(c, PK_E_STATE, "pk_ens_save: write failed");
    }
    return PK_OK;
}

int pk_ens_load(pk_ens* e, const char* path)
{
    if (!e || !path) return PK_E_ARG;
    pk_ctx* c = e->ctx;
    File in(path, "rb");
    PkSnapHeader h;
    if (int r = read_header(c, in, SNAP_ENSEMBLE, h)) return r;
    bool same = h.prec == (uint32_t)e->prec && h.posmode == (uint32_t)e->posmode && h.ghost == e->type &&
                h.nfield == e->nfield() && h.esize == (uint32_t)e->esize() && h.nsct == e->nsct && h.count >= 0;
    for (int d = 0; d < 3; d++)
        same = same && h.dims[d] == e->ncell[d] && h.ncps[d] == e->ncps[d] && h.step[d] == e->step[d];
    if (!same) return pk_fail(c, PK_E_ARG, "pk_ens_load: the snapshot's ensemble geometry, species or precision differs");
    const long nsct = e->nsct;
    std::vector<int32_t> cnt(nsct);
    if (!in.read(cnt.data(), sizeof(int32_t) * nsct)) return pk_fail(c, PK_E_ARG, "pk_ens_load: truncated file");
    // capacities as the re-sort sets them (count + count * frac + add)
    std::vector<int64_t> off(nsct + 1);
    int64_t np = 0, total = 0;
    int maxcap = 0;
    for (long s = 0; s < nsct; s++) {
        if (cnt[s] < 0) return pk_fail(c, PK_E_ARG, "pk_ens_load: negative supercell count");
        const long cap = (long)cnt[s] + (long)(cnt[s] * e->slack_frac) + e->slack_add;
        off[s] = total;
        total += cap;
        np += cnt[s];
        if (cap > maxcap) maxcap = (int)(cap < 0x7fffffff ? cap : 0x7fffffff);
    }
    off[nsct] = total;
    if (np != h.count) return pk_fail(c, PK_E_ARG, "pk_ens_load: supercell counts do not add up");
    if (maxcap > PK_MAX_SC_PARTICLES)
        return pk_fail(c, PK_E_CAPACITY, "pk_ens_load: supercell holds more than 65536 particles");
    if (int r = pk_ens_alloc(e, (size_t)total)) return r;
    std::vector<char> packed((size_t)np * h.esize), full((size_t)total * h.esize, 0);
    for (int f = 0; f < h.nfield; f++) {
        if (!in.read(packed.data(), packed.size())) return pk_fail(c, PK_E_ARG, "pk_ens_load: truncated file");
        size_t o = 0;
        for (long s = 0; s < nsct; s++) {
            std::memcpy(full.data() + (size_t)off[s] * h.esize, packed.data() + o, (size_t)cnt[s] * h.esize);
            o += (size_t)cnt[s] * h.esize;
        }
        PK_HIP(c, hipMemcpy(e->field(e->main, f), full.data(), full.size(), hipMemcpyHostToDevice));
    }
    PK_HIP(c, hipMemcpy(e->off, off.data(), sizeof(int64_t) * (nsct + 1), hipMemcpyHostToDevice));
    PK_HIP(c, hipMemcpy(e->cnt, cnt.data(), sizeof(int32_t) * nsct, hipMemcpyHostToDevice));
    PK_HIP(c, hipMemcpy(e->nstay, cnt.data(), sizeof(int32_t) * nsct, hipMemcpyHostToDevice));
    PK_HIP(c, hipMemset(e->nmove, 0, sizeof(int32_t) * nsct));
    PK_HIP(c, hipMemset(e->flags, 0, sizeof(int32_t) * PK_NFLAGS));
    e->maxcap = maxcap;
    e->np = np;
    e->pending_movers = false;
    e->failed = false;
    e->resorts = h.resorts;
    return PK_OK;
}

}  // extern "C"
