// blz_blastdb.cpp -- reader of BLAST protein databases as makeblastdb writes them (format
// versions 4 and 5), for loading a real database resident in HBM.
//
// The reference reads its database through NCBI's CSeqDB (BLAZE_ncbi/seqdbimpl.cpp:716-730, the
// sequence fetch its GPU path gathers from, gpuBLAZE.cu:1587-1618) and requires the FASTA to be
// sorted by length before makeblastdb (README.md:138).  Here the volume files are memory-mapped
// and gathered into the engine's layout (NCBIstdaa residues, u64 offsets) in OID order or in
// ascending length order with the OID permutation, so an unsorted database needs no rebuild.
//
// Layout (NCBI's published BLAST database format): the index file <name>.pin holds, big-endian
// unless noted: int32 format version (4 or 5); int32 sequence type (1 = protein); version 5 only:
// int32 volume number; int32 length + title; version 5 only: int32 length + LMDB file name;
// int32 length + date; int32 number of sequences N; int64 residue total (little-endian); int32
// longest sequence; N+1 int32 header offsets into <name>.phr; N+1 int32 sequence offsets into
// <name>.psq.  The sequence file starts with a NUL byte and every sequence is followed by one, so
// sequence i is psq[off[i], off[i+1] - 1).  An alias file <name>.pal lists volumes on a DBLIST line.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <memory>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "blz_common.h"
#include "blz_host.h"
#include "blz_prelim.h"

namespace {

thread_local std::string g_err;

blz_status fail(const std::string &msg) {
    g_err = msg;
    return BLZ_EINVAL;
}

struct Mapped {
    const uint8_t *p = nullptr;
    size_t n = 0;
    bool open(const std::string &path) {
        const int fd = ::open(path.c_str(), O_RDONLY);
        if (fd < 0) return false;
        struct stat sb;
        if (fstat(fd, &sb) != 0) {
            ::close(fd);
            return false;
        }
        n = (size_t)sb.st_size;
        if (n > 0) {
            void *m = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
            if (m == MAP_FAILED) {
                ::close(fd);
                return false;
            }
            p = static_cast<const uint8_t *>(m);
        }
        ::close(fd);
        return true;
    }
    void close() {
        if (p) munmap(const_cast<uint8_t *>(p), n);
        p = nullptr;
        n = 0;
    }
};

struct Volume {
    Mapped pin, psq, phr;
    int32_t version = 0;
    int64_t nseq = 0;
    uint64_t residues = 0;
    int32_t max_len = 0;
    std::string title, date;
    const uint8_t *hdr_off = nullptr;  // N+1 big-endian int32
    const uint8_t *seq_off = nullptr;  // N+1 big-endian int32
};

uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }
uint64_t le64(const uint8_t *p) {
    uint64_t v = 0;
    for (int i = 7; i >= 0; i--) v = v << 8 | p[i];
    return v;
}

blz_status parse_index(Volume &v, const std::string &name) {
    const uint8_t *p = v.pin.p;
    const size_t n = v.pin.n;
    size_t at = 0;
    auto need = [&](size_t k) { return at + k <= n; };
    auto int32 = [&](uint32_t &x) {
        if (!need(4)) return false;
        x = be32(p + at);
        at += 4;
        return true;
    };
    auto str = [&](std::string &s) {
        uint32_t len;
        if (!int32(len) || !need(len)) return false;
        s.assign(reinterpret_cast<const char *>(p + at), len);
        at += len;
        return true;
    };
    uint32_t version, type, u;
    std::string lmdb;
    if (!int32(version) || (version != 4 && version != 5)) return fail(name + ".pin: format version is not 4 or 5");
    if (!int32(type) || type != 1) return fail(name + ".pin: not a protein database");
    if (version == 5 && !int32(u)) return fail(name + ".pin: truncated");
    if (!str(v.title) || (version == 5 && !str(lmdb)) || !str(v.date)) return fail(name + ".pin: truncated header");
    uint32_t nseq, maxlen;
    if (!int32(nseq) || !need(8)) return fail(name + ".pin: truncated header");
    v.residues = le64(p + at);
    at += 8;
    if (!int32(maxlen)) return fail(name + ".pin: truncated header");
    if (!need(2 * ((size_t)nseq + 1) * 4)) return fail(name + ".pin: offset tables truncated");
    v.version = (int32_t)version;
    v.nseq = nseq;
    v.max_len = (int32_t)maxlen;
    v.hdr_off = p + at;
    v.seq_off = p + at + ((size_t)nseq + 1) * 4;
    // the offset tables must rise within their files, and the residue total and longest sequence
    // must agree with them
    uint64_t total = 0;
    uint32_t longest = 0;
    for (uint32_t i = 0; i < nseq; i++) {
        const uint32_t s0 = be32(v.seq_off + 4 * i), s1 = be32(v.seq_off + 4 * (i + 1));
        const uint32_t h0 = be32(v.hdr_off + 4 * i), h1 = be32(v.hdr_off + 4 * (i + 1));
        if (s1 <= s0 || s1 > v.psq.n || h1 < h0 || h1 > v.phr.n)
            return fail(name + ": sequence or header offsets outside their files");
        total += s1 - s0 - 1;
        longest = std::max(longest, s1 - s0 - 1);
    }
    if (nseq > 0 && (be32(v.seq_off) != 1 || v.psq.p[0] != 0))
        return fail(name + ".psq: does not start with the NUL separator");
    if (total != v.residues) return fail(name + ".pin: residue total does not match the sequence offsets");
    if (nseq > 0 && longest != maxlen) return fail(name + ".pin: longest-sequence field does not match");
    return BLZ_OK;
}

// Volume base paths of a database name: an alias file's DBLIST (relative to its directory), or
// the name itself.
bool volume_names(const std::string &path, std::vector<std::string> &out) {
    std::ifstream pal(path + ".pal");
    if (!pal) {
        out.push_back(path);
        return true;
    }
    const size_t slash = path.find_last_of('/');
    const std::string dir = slash == std::string::npos ? "" : path.substr(0, slash + 1);
    std::string line;
    while (std::getline(pal, line)) {
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key) || key != "DBLIST") continue;
        std::string vol;
        while (ls >> vol) {
            if (vol.size() >= 2 && vol.front() == '"' && vol.back() == '"') vol = vol.substr(1, vol.size() - 2);
            out.push_back(!vol.empty() && vol[0] == '/' ? vol : dir + vol);
        }
    }
    return !out.empty();
}

}  // namespace

struct blz_blastdb {
    std::vector<Volume> vols;
    std::vector<int64_t> first;  // first OID of each volume, and the total
    uint64_t residues = 0;
    int32_t max_len = 0;
    std::string title;
};

extern "C" {

const char *blz_blastdb_last_error(void) { return g_err.c_str(); }

void blz_blastdb_close(blz_blastdb *db) {
    if (!db) return;
    for (Volume &v : db->vols) {
        v.pin.close();
        v.psq.close();
        v.phr.close();
    }
    delete db;
}

blz_status blz_blastdb_open(const char *path, blz_blastdb **out) {
    if (!path || !out) return BLZ_EINVAL;
    *out = nullptr;
    std::vector<std::string> names;
    if (!volume_names(path, names)) return fail(std::string(path) + ".pal: no DBLIST line");
    blz_blastdb *db = new blz_blastdb();
    db->vols.resize(names.size());
    db->first.push_back(0);
    for (size_t k = 0; k < names.size(); k++) {
        Volume &v = db->vols[k];
        if (!v.pin.open(names[k] + ".pin") || !v.psq.open(names[k] + ".psq") || !v.phr.open(names[k] + ".phr")) {
            blz_blastdb_close(db);
            return fail(names[k] + ": cannot open .pin, .psq and .phr");
        }
        const blz_status st = parse_index(v, names[k]);
        if (st != BLZ_OK) {
            blz_blastdb_close(db);
            return st;
        }
        db->first.push_back(db->first.back() + v.nseq);
        db->residues += v.residues;
        db->max_len = std::max(db->max_len, v.max_len);
    }
    db->title = db->vols.size() == 1 ? db->vols[0].title : std::string(path);
    *out = db;
    return BLZ_OK;
}

int64_t blz_blastdb_num_seqs(const blz_blastdb *db) { return db ? db->first.back() : 0; }
uint64_t blz_blastdb_num_residues(const blz_blastdb *db) { return db ? db->residues : 0; }
int32_t blz_blastdb_max_length(const blz_blastdb *db) { return db ? db->max_len : 0; }
const char *blz_blastdb_title(const blz_blastdb *db) { return db ? db->title.c_str() : ""; }

// The volume holding an OID and the OID's index in it.
static const Volume *locate(const blz_blastdb *db, int64_t oid, int64_t &local) {
    const auto it = std::upper_bound(db->first.begin(), db->first.end(), oid);
    const size_t k = (size_t)(it - db->first.begin()) - 1;
    local = oid - db->first[k];
    return &db->vols[k];
}

blz_status blz_blastdb_sequence(const blz_blastdb *db, int64_t oid, const uint8_t **letters, int32_t *len) {
    if (!db || !letters || !len || oid < 0 || oid >= db->first.back()) return BLZ_EINVAL;
    int64_t i;
    const Volume *v = locate(db, oid, i);
    const uint32_t s0 = be32(v->seq_off + 4 * i), s1 = be32(v->seq_off + 4 * (i + 1));
    *letters = v->psq.p + s0;
    *len = (int32_t)(s1 - s0 - 1);
    return BLZ_OK;
}

blz_status blz_blastdb_header(const blz_blastdb *db, int64_t oid, const uint8_t **bytes, int64_t *len) {
    if (!db || !bytes || !len || oid < 0 || oid >= db->first.back()) return BLZ_EINVAL;
    int64_t i;
    const Volume *v = locate(db, oid, i);
    const uint32_t h0 = be32(v->hdr_off + 4 * i), h1 = be32(v->hdr_off + 4 * (i + 1));
    *bytes = v->phr.p + h0;
    *len = (int64_t)(h1 - h0);
    return BLZ_OK;
}

blz_status blz_blastdb_gather(const blz_blastdb *db, int32_t sort_by_length, uint8_t *residues, uint64_t *offsets,
                              int64_t *perm) {
    if (!db || !residues || !offsets) return BLZ_EINVAL;
    const int64_t n = db->first.back();
    std::vector<int64_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::vector<int32_t> lens(n);
    for (int64_t oid = 0; oid < n; oid++) {
        const uint8_t *s;
        blz_blastdb_sequence(db, oid, &s, &lens[oid]);
    }
    // ascending length, OIDs in order within a length (the order of a length-sorted FASTA)
    if (sort_by_length)
        std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return lens[a] < lens[b]; });
    offsets[0] = 0;
    bool bad = false;
    for (int64_t i = 0; i < n; i++) {
        const uint8_t *s;
        int32_t len;
        blz_blastdb_sequence(db, order[i], &s, &len);
        uint8_t any = 0;
        for (int32_t k = 0; k < len; k++) any |= (uint8_t)(s[k] >= blz::kAlpha);
        bad |= any != 0;
        std::memcpy(residues + offsets[i], s, (size_t)len);
        offsets[i + 1] = offsets[i] + (uint64_t)len;
        if (perm) perm[i] = order[i];
    }
    return bad ? fail("sequence letter outside NCBIstdaa (>= 28)") : BLZ_OK;
}

blz_status blz_db_upload_blastdb(blz_ctx *ctx, const blz_blastdb *bdb, int32_t sort_by_length, blz_db **out,
                                 int64_t *perm) {
    if (!ctx || !bdb || !out) return BLZ_EINVAL;
    const int64_t n = bdb->first.back();
    std::vector<uint8_t> res(bdb->residues + 1);
    std::vector<uint64_t> offs((size_t)n + 1);
    auto ids = std::make_shared<std::vector<int64_t>>((size_t)std::max<int64_t>(n, 1));
    blz_status st = blz_blastdb_gather(bdb, sort_by_length, res.data(), offs.data(), ids->data());
    if (st != BLZ_OK) return st;
    st = blz_db_upload(ctx, res.data(), offs.data(), n, out);
    if (st != BLZ_OK) return st;
    ids->resize((size_t)n);
    if (perm) std::copy(ids->begin(), ids->end(), perm);
    if (sort_by_length) (*out)->oid_of = std::move(ids);  // the hit list breaks ties by these ids
    return BLZ_OK;
}

}  // extern "C"
