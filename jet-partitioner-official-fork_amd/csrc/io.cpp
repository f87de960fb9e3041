// File formats of the reference's drivers (app/io.hpp:50-244): the METIS
// graph format, the config file and one-part-per-line partition files.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

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
int jet_load_metis(const char* path, jet_host_graph_t* out) {
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
            while (f < fmax && std::isdigit((unsigned char)*f)) v = v * 10 + (uint64_t)(*f++ - '0');
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
        while (f < fmax && std::isdigit((unsigned char)*f)) v = v * 10 + (*f++ - '0');
        if (!is_value) {
            if (v < 1 || v > n) return fail(JET_EINVAL, "metis neighbour id out of range");
            col.push_back((int32_t)(v - 1));
        } else {
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

}  // extern "C"
