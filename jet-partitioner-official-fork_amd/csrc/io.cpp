// File formats of the reference's drivers (app/io.hpp:50-244): the METIS
// graph format, the config file and one-part-per-line partition files.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "hierarchy.h"
#include "jet_capi.h"

namespace jet {
void set_last_error(const std::string& msg);
}

namespace {

int fail(int code, const std::string& msg) {
    jet::set_last_error(msg);
    return code;
}

// Reads a whole file; appends a newline when the last line lacks one.
bool slurp(const char* path, std::vector<char>& buf) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) return false;
    f.seekg(0, std::ios::end);
    std::streamoff sz = f.tellg();
    f.seekg(0, std::ios::beg);
    buf.resize((size_t)sz + 1);
    if (sz > 0) f.read(buf.data(), sz);
    if (sz == 0 || buf[(size_t)sz - 1] != '\n') buf[(size_t)sz++] = '\n';
    buf.resize((size_t)sz);
    return true;
}

}  // namespace

extern "C" {

// app/io.hpp:50-81: at least four whitespace-separated values (coarsening
// algorithm, parts, attempts, imbalance) and an optional fifth (ultra).
int jet_load_config(const char* path, jet_config_t* cfg) {
    if (!path || !cfg) return fail(JET_EINVAL, "path and cfg are required");
    std::ifstream f(path);
    if (!f.is_open()) return fail(JET_EINVAL, std::string("could not open config file ") + path);
    std::string tok[5];
    int reads = 0;
    for (int i = 0; i < 5; i++)
        if (f >> tok[i]) reads++;
    if (reads < 4) return fail(JET_EINVAL, "config file has fewer than 4 values");
    jet_config_default(cfg);
    try {
        cfg->coarsening_alg = std::stoi(tok[0]);
        cfg->num_parts = std::stoi(tok[1]);
        cfg->num_iter = std::stoi(tok[2]);
        cfg->max_imb_ratio = std::stod(tok[3]);
        if (reads == 5) cfg->ultra_settings = std::stoi(tok[4]);
    } catch (const std::exception&) {
        return fail(JET_EINVAL, "config file value is not a number");
    }
    return JET_OK;
}

// app/io.hpp:96-205: header "n m [fmt [ncon]]" after '%' comment lines, then
// one line of 1-based neighbours per vertex, with a weight after each
// neighbour when fmt is 1. Vertex weights (fmt 10/11) and ncon are not
// supported, as in the reference.
static int load_metis(const char* path, jet_host_graph_t* out) {
    if (!path || !out) return fail(JET_EINVAL, "path and out are required");
    std::memset(out, 0, sizeof(*out));
    std::vector<char> buf;
    if (!slurp(path, buf)) return fail(JET_EINVAL, std::string("could not open metis graph file ") + path);
    const char* f = buf.data();
    const char* fmax = f + buf.size();
    auto next_line = [&]() {
        while (f < fmax && *f != '\n') f++;
        if (f < fmax) f++;
    };
    while (f < fmax && *f == '%') next_line();
    uint64_t header[4] = {0, 0, 0, 0};
    int nh = 0;
    while (f < fmax && *f != '\n') {
        if (std::isdigit((unsigned char)*f)) {
            uint64_t v = 0;
            bool big = false;
            while (f < fmax && std::isdigit((unsigned char)*f)) {
                if (v > (UINT64_MAX - 9) / 10) big = true;
                else v = v * 10 + (uint64_t)(*f - '0');
                f++;
            }
            if (big) return fail(JET_EINVAL, "metis header value out of range");
            if (nh < 4) header[nh] = v;
            nh++;
        } else {
            f++;
        }
    }
    if (f < fmax) f++;
    if (nh < 2) return fail(JET_EINVAL, "metis header needs at least n and m");
    const int64_t n = (int64_t)header[0], m = (int64_t)header[1];
    const int fmt = (int)header[2], ncon = (int)header[3];
    if (fmt != 0 && fmt != 1) return fail(JET_EUNSUPPORTED, "unsupported metis format flags (vertex weights)");
    if (ncon != 0) return fail(JET_EUNSUPPORTED, "unsupported metis ncon");
    if (n < 1 || n > 2147483645LL) return fail(JET_EINVAL, "metis n out of range");
    // every stored neighbour takes at least two characters of the file
    if (header[1] > (uint64_t)buf.size()) return fail(JET_EINVAL, "metis m exceeds what the file can hold");
    const bool has_ew = fmt == 1;
    std::vector<int64_t> rp((size_t)n + 1, 0);
    std::vector<int32_t> col, ew;
    col.reserve((size_t)(2 * m));
    if (has_ew) ew.reserve((size_t)(2 * m));
    int64_t rows = 0;
    bool is_value = false;
    while (f < fmax) {
        while (f < fmax && !std::isdigit((unsigned char)*f)) {
            if (*f == '%') {
                next_line();
                continue;
            }
            if (*f == '\n' && rows < n) rp[(size_t)++rows] = (int64_t)col.size();
            f++;
        }
        if (f >= fmax) break;
        int64_t v = 0;
        while (f < fmax && std::isdigit((unsigned char)*f)) {
            if (v <= INT32_MAX) v = v * 10 + (*f - '0');  // saturates above INT32_MAX * 10
            f++;
        }
        if (!is_value) {
            if (v < 1 || v > n) return fail(JET_EINVAL, "metis neighbour id out of range");
            col.push_back((int32_t)(v - 1));
        } else {
            if (v > INT32_MAX) return fail(JET_EINVAL, "metis edge weight out of range");
            ew.push_back((int32_t)v);
        }
        if (has_ew) is_value = !is_value;
    }
    if (rows != n || (int64_t)col.size() != 2 * m || (has_ew && ew.size() != col.size()))
        return fail(JET_EINVAL, "metis line or nonzero count does not match the header (read " + std::to_string(rows) +
                                    " lines, " + std::to_string(col.size()) + " nonzeros)");
    out->n = n;
    out->m = (int64_t)col.size();
    out->row_offsets = static_cast<int64_t*>(std::malloc(sizeof(int64_t) * (size_t)(n + 1)));
    out->col = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * std::max<size_t>(col.size(), 1)));
    out->ew = has_ew ? static_cast<int32_t*>(std::malloc(sizeof(int32_t) * std::max<size_t>(ew.size(), 1))) : nullptr;
    if (!out->row_offsets || !out->col || (has_ew && !out->ew)) {
        jet_free_graph(out);
        return fail(JET_ENOMEM, "host allocation failed");
    }
    std::memcpy(out->row_offsets, rp.data(), sizeof(int64_t) * rp.size());
    std::memcpy(out->col, col.data(), sizeof(int32_t) * col.size());
    if (has_ew) std::memcpy(out->ew, ew.data(), sizeof(int32_t) * ew.size());
    out->uniform_ew = has_ew ? 0 : 1;
    return JET_OK;
}

int jet_load_metis(const char* path, jet_host_graph_t* out) {
    try {
        return load_metis(path, out);
    } catch (const std::bad_alloc&) {
        if (out) jet_free_graph(out);
        return fail(JET_ENOMEM, "host allocation failed while reading the metis file");
    } catch (const std::exception& e) {
        if (out) jet_free_graph(out);
        return fail(JET_EINVAL, std::string("metis file: ") + e.what());
    }
}

void jet_free_graph(jet_host_graph_t* g) {
    if (!g) return;
    std::free(g->row_offsets);
    std::free(g->col);
    std::free(g->ew);
    std::memset(g, 0, sizeof(*g));
}

// app/io.hpp:207-220: one part id per line.
int jet_write_partition(const char* path, int64_t n, const int32_t* part) {
    if (!path || (!part && n > 0)) return fail(JET_EINVAL, "path and part are required");
    FILE* fp = std::fopen(path, "w");
    if (!fp) return fail(JET_EINVAL, std::string("could not open ") + path + " for writing");
    for (int64_t i = 0; i < n; i++) std::fprintf(fp, "%d\n", part[i]);
    bool ok = std::fclose(fp) == 0;
    return ok ? JET_OK : fail(JET_EINVAL, "write failed");
}

// app/io.hpp:222-235: n part ids, whitespace separated.
int jet_read_partition(const char* path, int64_t n, int32_t* part) {
    if (!path || (!part && n > 0)) return fail(JET_EINVAL, "path and part are required");
    std::ifstream f(path);
    if (!f.is_open()) return fail(JET_EINVAL, std::string("could not open partition file ") + path);
    for (int64_t i = 0; i < n; i++)
        if (!(f >> part[i])) return fail(JET_EINVAL, "partition file has fewer than n entries");
    return JET_OK;
}

// import_coarse.cpp:49-110: the hierarchy file. Values are read at the widths
// of the reference graph type that wrote them and checked: sizes against the
// file, offsets monotone from 0, columns and map entries in range, weights in
// [1, 2^31).
int jet_load_hierarchy(const char* path, int32_t kind, jet_hierarchy_t** out) {
    if (!path || !out) return fail(JET_EINVAL, "path and out are required");
    *out = nullptr;
    if (kind != JET_DUMP_32 && kind != JET_DUMP_BIG && kind != JET_DUMP_BIGGEST)
        return fail(JET_EINVAL, "kind must be JET_DUMP_32, JET_DUMP_BIG or JET_DUMP_BIGGEST");
    const int ob = kind == JET_DUMP_BIGGEST ? 8 : 4;  // ordinal bytes
    const int eb = kind == JET_DUMP_32 ? 4 : 8;       // offset and weight bytes
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(JET_EINVAL, std::string("could not open hierarchy file ") + path);
    auto rd = [&](int bytes, int64_t count, std::vector<int64_t>& v) -> bool {
        v.resize((size_t)count);
        if (bytes == 8) return count == 0 || std::fread(v.data(), 8, (size_t)count, f) == (size_t)count;
        std::vector<int32_t> t((size_t)count);
        if (count > 0 && std::fread(t.data(), 4, (size_t)count, f) != (size_t)count) return false;
        for (int64_t i = 0; i < count; i++) v[(size_t)i] = t[(size_t)i];
        return true;
    };
    auto bad = [&](const std::string& msg) {
        std::fclose(f);
        return fail(JET_EINVAL, "hierarchy file " + std::string(path) + ": " + msg);
    };
    int32_t count = 0;
    if (std::fread(&count, 4, 1, f) != 1 || count < 1) return bad("missing or non-positive level count");
    std::unique_ptr<jet_hierarchy_t> h(new jet_hierarchy_t());
    std::vector<int64_t> t;
    int64_t prev_n = 0;
    for (int32_t l = 0; l < count; l++) {
        jet_hierarchy::Level L;
        std::vector<int64_t> hdr;
        if (!rd(ob, 1, hdr)) return bad("truncated at level " + std::to_string(l));
        L.n = hdr[0];
        if (!rd(eb, 1, hdr)) return bad("truncated at level " + std::to_string(l));
        L.m = hdr[0];
        if (L.n < 1 || L.n > 2147483645LL || L.m < 0) return bad("level " + std::to_string(l) + ": bad n or m");
        if (!rd(eb, L.n + 1, L.rowptr)) return bad("truncated row offsets at level " + std::to_string(l));
        if (L.rowptr[0] != 0 || L.rowptr[(size_t)L.n] != L.m) return bad("row offsets do not span m");
        for (int64_t i = 0; i < L.n; i++)
            if (L.rowptr[(size_t)i + 1] < L.rowptr[(size_t)i]) return bad("row offsets not monotone");
        auto narrow = [&](int bytes, int64_t cnt, std::vector<int32_t>& dst, int64_t lo, int64_t hi) -> int {
            if (!rd(bytes, cnt, t)) return 1;
            dst.resize((size_t)cnt);
            for (int64_t i = 0; i < cnt; i++) {
                const int64_t v = t[(size_t)i];
                if (v < lo || v > hi) return 2;
                dst[(size_t)i] = (int32_t)v;
            }
            return 0;
        };
        int e = narrow(ob, L.m, L.col, 0, L.n - 1);
        if (e) return bad(e == 1 ? "truncated columns" : "column out of range");
        e = narrow(eb, L.m, L.ew, 1, 2147483647LL);
        if (e) return bad(e == 1 ? "truncated edge weights" : "edge weight outside [1, 2^31)");
        e = narrow(eb, L.n, L.vw, 1, 2147483647LL);
        if (e) return bad(e == 1 ? "truncated vertex weights" : "vertex weight outside [1, 2^31)");
        if (l > 0) {
            e = narrow(ob, prev_n, L.map, 0, L.n - 1);
            if (e) return bad(e == 1 ? "truncated map" : "map entry out of range");
        }
        prev_n = L.n;
        h->levels.push_back(std::move(L));
    }
    std::fclose(f);
    *out = h.release();
    return JET_OK;
}

int32_t jet_hierarchy_levels(const jet_hierarchy_t* h) { return h ? (int32_t)h->levels.size() : 0; }

int64_t jet_hierarchy_vertices(const jet_hierarchy_t* h, int32_t level) {
    if (!h || level < 0 || level >= (int32_t)h->levels.size()) return -1;
    return h->levels[(size_t)level].n;
}

int64_t jet_hierarchy_edges(const jet_hierarchy_t* h, int32_t level) {
    if (!h || level < 0 || level >= (int32_t)h->levels.size()) return -1;
    return h->levels[(size_t)level].m;
}

void jet_free_hierarchy(jet_hierarchy_t* h) { delete h; }

// import_coarse.cpp:112-121 and binary_dump.hpp:89-95: n raw int32 part ids.
int jet_load_coarse_part(const char* path, int64_t n, int32_t* part) {
    if (!path || !part || n < 1) return fail(JET_EINVAL, "path, part and n >= 1 are required");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(JET_EINVAL, std::string("could not open coarse partition file ") + path);
    const size_t got = std::fread(part, 4, (size_t)n, f);
    std::fclose(f);
    if (got != (size_t)n) return fail(JET_EINVAL, "coarse partition file has fewer than n entries");
    return JET_OK;
}

}  // extern "C"
