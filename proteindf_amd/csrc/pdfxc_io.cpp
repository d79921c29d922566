// ProteinDF's dense matrix files and the grid matrix (grid.mat), on the host.
//
// File layout (TlMatrixObject, tl_matrix_object.h:50-61): a header of the
// matrix type (RSFD 0, CSFD 1, RLHD 2, ...), the row count and the column
// count, then the doubles.  The reference writes the type as one byte and the
// counts as 32-bit ints (tl_dense_symmetric_matrix_object.cc:292-312,
// tl_dense_general_matrix_object.cc:239-266); on reading it accepts any of
// the widths (type 1, 4 or 8 bytes; counts 4 or 8 bytes) and takes the first,
// in that order, whose implied file size is the real one
// (tl_matrix_utils.cc:114-200 with estimateFileSize :334-358).
//   RSFD: row-major, rows x cols.   CSFD: column-major, rows x cols.
//   RLHD: symmetric, rows == cols, lower triangle by rows: (0,0) (1,0) (1,1) ...
// A symmetric matrix loads from RLHD only (tl_dense_symmetric_matrix_object.cc
// :245-289); a general matrix loads from RSFD or CSFD and saves as CSFD.
//
// The grid matrix (DfGenerateGrid.cpp:229-239, :748-767; DfCalcGridX.h:37-54)
// is a general matrix with one row per point: x, y, z, weight, atom, then
// rho (LDA RKS) or rho, d/dx, d/dy, d/dz (GGA RKS).
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pdfxc_internal.hpp"

namespace pdfxc {

enum { kRSFD = 0, kCSFD = 1, kRLHD = 2, kRUHD = 3, kCLHD = 4, kCUHD = 5 };

struct MatFile {
    int type = -1;
    int64_t rows = 0, cols = 0;
    int64_t header = 0;  // bytes before the first double
};

static int io_error(const std::string& msg) {
    set_error(msg);
    return PDFXC_ERR_ARGUMENT;
}

static int64_t le_signed(const unsigned char* p, int width) {
    uint64_t v = 0;
    for (int k = 0; k < width; ++k) v |= (uint64_t)p[k] << (8 * k);
    if (width < 8 && (v >> (8 * width - 1)) & 1) v |= ~(uint64_t)0 << (8 * width);  // sign-extend
    return (int64_t)v;
}

// number of doubles the body holds for (type, rows, cols), -1 if inconsistent
static int64_t body_items(int type, int64_t rows, int64_t cols) {
    if (rows < 0 || cols < 0 || rows > INT32_MAX || cols > INT32_MAX) return -1;
    switch (type) {
        case kRSFD:
        case kCSFD: return rows * cols;
        case kRLHD:
        case kRUHD:
        case kCLHD:
        case kCUHD: return rows == cols ? rows * (rows + 1) / 2 : -1;
        default: return -1;
    }
}

static int open_matrix(const char* path, FILE** out, MatFile* h) {
    if (!path) return io_error("matrix file: null path");
    FILE* f = std::fopen(path, "rb");
    if (!f) return io_error(std::string("cannot open matrix file: ") + path);
    std::fseek(f, 0, SEEK_END);
    const int64_t size = (int64_t)std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    unsigned char buf[24] = {0};
    const size_t got = std::fread(buf, 1, sizeof(buf), f);
    static const int kWidths[6][2] = {{1, 4}, {1, 8}, {4, 4}, {4, 8}, {8, 4}, {8, 8}};
    for (const auto& wd : kWidths) {
        const int tw = wd[0], iw = wd[1];
        if ((int64_t)got < tw + 2 * iw) continue;
        const int64_t type = le_signed(buf, tw);
        const int64_t rows = le_signed(buf + tw, iw), cols = le_signed(buf + tw + iw, iw);
        if (type < kRSFD || type > kCUHD) continue;
        const int64_t items = body_items((int)type, rows, cols);
        if (items < 0 || tw + 2 * iw + 8 * items != size) continue;
        h->type = (int)type;
        h->rows = rows;
        h->cols = cols;
        h->header = tw + 2 * iw;
        std::fseek(f, (long)h->header, SEEK_SET);
        *out = f;
        return PDFXC_OK;
    }
    std::fclose(f);
    set_error(std::string("illegal matrix format (no dense header matches the file size): ") + path);
    return PDFXC_ERR_UNSUPPORTED;
}

static int read_doubles(FILE* f, double* dst, int64_t n, const char* path) {
    const int64_t chunk = 1 << 20;
    for (int64_t done = 0; done < n;) {
        const int64_t k = std::min(chunk, n - done);
        if ((int64_t)std::fread(dst + done, sizeof(double), (size_t)k, f) != k)
            return io_error(std::string("short read: ") + path);
        done += k;
    }
    return PDFXC_OK;
}

static int write_matrix(const char* path, int type, int64_t rows, int64_t cols, const double* data,
                        int64_t n) {
    if (!path) return io_error("matrix file: null path");
    if (rows > INT32_MAX || cols > INT32_MAX) return io_error("matrix file: dimension beyond 32 bits");
    FILE* f = std::fopen(path, "wb");
    if (!f) return io_error(std::string("cannot write matrix: ") + path);
    const char t = (char)type;
    const int32_t r = (int32_t)rows, c = (int32_t)cols;
    bool ok = std::fwrite(&t, 1, 1, f) == 1 && std::fwrite(&r, 4, 1, f) == 1 && std::fwrite(&c, 4, 1, f) == 1;
    if (ok && n > 0) ok = (int64_t)std::fwrite(data, sizeof(double), (size_t)n, f) == n;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? PDFXC_OK : io_error(std::string("write failed: ") + path);
}

// caller-order rho (n_pts x 4) <-> the context's slots
static int download_rho(Context* c, double* rho, const double* d_src) {
    std::vector<double> rs(4 * (size_t)c->n_slots);
    PDFXC_HIP(hipStreamSynchronize(c->stream));
    PDFXC_HIP(hipMemcpy(rs.data(), d_src, rs.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t p = 0; p < c->n_pts; ++p)
        for (int q = 0; q < 4; ++q) rho[4 * p + q] = rs[4 * c->slot_of_point[p] + q];
    return PDFXC_OK;
}

static int upload_rho(Context* c, const double* rho, double* d_dst) {
    std::vector<double> rs(4 * (size_t)c->n_slots, 0.0);
    for (int64_t p = 0; p < c->n_pts; ++p)
        for (int q = 0; q < 4; ++q) rs[4 * c->slot_of_point[p] + q] = rho[4 * p + q];
    PDFXC_HIP(hipStreamSynchronize(c->stream));
    PDFXC_HIP(hipMemcpy(d_dst, rs.data(), rs.size() * sizeof(double), hipMemcpyHostToDevice));
    return PDFXC_OK;
}

}  // namespace pdfxc

using namespace pdfxc;

extern "C" {

int pdfxc_matrix_file_info(const char* path, int* type, int64_t* rows, int64_t* cols) {
    FILE* f = nullptr;
    MatFile h;
    if (int rc = open_matrix(path, &f, &h)) return rc;
    std::fclose(f);
    if (type) *type = h.type;
    if (rows) *rows = h.rows;
    if (cols) *cols = h.cols;
    return PDFXC_OK;
}

int pdfxc_load_symmetric_matrix(const char* path, int64_t n, double* packed) {
    if (n < 0 || (n > 0 && !packed)) return io_error("pdfxc_load_symmetric_matrix: bad arguments");
    FILE* f = nullptr;
    MatFile h;
    if (int rc = open_matrix(path, &f, &h)) return rc;
    int rc = PDFXC_OK;
    if (h.type != kRLHD) {
        set_error(std::string("not supported format for a symmetric matrix (RLHD only): ") + path);
        rc = PDFXC_ERR_UNSUPPORTED;
    } else if (h.rows != n) {
        rc = io_error("pdfxc_load_symmetric_matrix: file holds dimension " + std::to_string(h.rows) + ", expected " +
                      std::to_string(n) + ": " + path);
    } else {
        rc = read_doubles(f, packed, n * (n + 1) / 2, path);
    }
    std::fclose(f);
    return rc;
}

int pdfxc_save_symmetric_matrix(const char* path, int64_t n, const double* packed) {
    if (n < 0 || (n > 0 && !packed)) return io_error("pdfxc_save_symmetric_matrix: bad arguments");
    return write_matrix(path, kRLHD, n, n, packed, n * (n + 1) / 2);
}

int pdfxc_load_general_matrix(const char* path, int64_t rows, int64_t cols, double* colmajor) {
    if (rows < 0 || cols < 0 || (rows * cols > 0 && !colmajor))
        return io_error("pdfxc_load_general_matrix: bad arguments");
    FILE* f = nullptr;
    MatFile h;
    if (int rc = open_matrix(path, &f, &h)) return rc;
    int rc = PDFXC_OK;
    if (h.type != kRSFD && h.type != kCSFD) {
        set_error(std::string("not supported format for a general matrix (RSFD or CSFD): ") + path);
        rc = PDFXC_ERR_UNSUPPORTED;
    } else if (h.rows != rows || h.cols != cols) {
        rc = io_error("pdfxc_load_general_matrix: file holds " + std::to_string(h.rows) + " x " +
                      std::to_string(h.cols) + ", expected " + std::to_string(rows) + " x " +
                      std::to_string(cols) + ": " + path);
    } else if (h.type == kCSFD) {
        rc = read_doubles(f, colmajor, rows * cols, path);
    } else {
        std::vector<double> row((size_t)cols);
        for (int64_t r = 0; r < rows && rc == PDFXC_OK; ++r) {
            rc = read_doubles(f, row.data(), cols, path);
            for (int64_t c = 0; c < cols; ++c) colmajor[r + rows * c] = row[c];
        }
    }
    std::fclose(f);
    return rc;
}

int pdfxc_save_general_matrix(const char* path, int64_t rows, int64_t cols, const double* colmajor) {
    if (rows < 0 || cols < 0 || (rows * cols > 0 && !colmajor))
        return io_error("pdfxc_save_general_matrix: bad arguments");
    return write_matrix(path, kCSFD, rows, cols, colmajor, rows * cols);
}

int pdfxc_load_grid_matrix(pdfxc_ctx* ctx, const char* path, int64_t* n_pts_out, int* n_cols_out) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return io_error("pdfxc_load_grid_matrix: null context");
    FILE* f = nullptr;
    MatFile h;
    if (int rc = open_matrix(path, &f, &h)) return rc;
    std::fclose(f);
    if ((h.type != kRSFD && h.type != kCSFD) || h.cols < 4 || h.rows < 1) {
        set_error(std::string("pdfxc_load_grid_matrix: not a grid matrix (general, >= 4 columns): ") + path);
        return PDFXC_ERR_UNSUPPORTED;
    }
    const int64_t n = h.rows, nc = h.cols;
    std::vector<double> m((size_t)(n * nc));
    if (int rc = pdfxc_load_general_matrix(path, n, nc, m.data())) return rc;
    const double* col = m.data();
    if (int rc = pdfxc_set_grid(ctx, n, col, col + n, col + 2 * n, col + 3 * n)) return rc;
    if (nc >= 5) {
        std::vector<int> atom((size_t)n);
        for (int64_t p = 0; p < n; ++p) {
            const double a = col[4 * n + p];
            if (!(a >= 0.0) || a != (double)(int)a)
                return io_error("pdfxc_load_grid_matrix: bad atom index at point " + std::to_string(p));
            atom[p] = (int)a;
        }
        c->gen_atom = std::move(atom);
    }
    // rho columns (6: LDA RKS, 7: LDA UKS, 9: GGA RKS, 13: GGA UKS): continue an
    // update-mode SCF from the file
    const int per_spin = (nc == 6 || nc == 7) ? 1 : (nc == 9 || nc == 13) ? 4 : 0;
    const int nspin = (nc == 7 || nc == 13) ? 2 : 1;
    for (int sp = 0; per_spin > 0 && sp < nspin; ++sp) {
        std::vector<double> rho(4 * (size_t)n, 0.0);
        for (int64_t p = 0; p < n; ++p)
            for (int q = 0; q < per_spin; ++q) rho[4 * p + q] = col[(5 + per_spin * sp + q) * n + p];
        if (sp == 1)
            if (int rc = ensure_beta_buffers(c)) return rc;
        if (int rc = upload_rho(c, rho.data(), sp ? c->d_rho_b : c->d_rho)) return rc;
    }
    if (n_pts_out) *n_pts_out = n;
    if (n_cols_out) *n_cols_out = (int)nc;
    return PDFXC_OK;
}

int pdfxc_save_grid_matrix(pdfxc_ctx* ctx, const char* path, int n_rho_cols) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return io_error("pdfxc_save_grid_matrix: null context");
    if (n_rho_cols != 0 && n_rho_cols != 1 && n_rho_cols != 2 && n_rho_cols != 4 && n_rho_cols != 8)
        return io_error("pdfxc_save_grid_matrix: n_rho_cols must be 0, 1 / 2 (LDA RKS / UKS) or 4 / 8 (GGA RKS / UKS)");
    if ((n_rho_cols == 2 || n_rho_cols == 8) && !c->d_rho_b) {
        set_error("pdfxc_save_grid_matrix: no UKS build has run on this grid");
        return PDFXC_ERR_ARGUMENT;
    }
    if (c->n_pts == 0) {
        set_error("pdfxc_save_grid_matrix: no grid");
        return PDFXC_ERR_NO_GRID;
    }
    const int64_t n = c->n_pts, nc = 5 + n_rho_cols;
    std::vector<double> m((size_t)(n * nc));
    std::copy(c->gx.begin(), c->gx.end(), m.begin());
    std::copy(c->gy.begin(), c->gy.end(), m.begin() + n);
    std::copy(c->gz.begin(), c->gz.end(), m.begin() + 2 * n);
    std::copy(c->gw.begin(), c->gw.end(), m.begin() + 3 * n);
    for (int64_t p = 0; p < n; ++p) m[4 * n + p] = c->gen_atom.size() == (size_t)n ? c->gen_atom[p] : -1.0;
    const int nspin = (n_rho_cols == 2 || n_rho_cols == 8) ? 2 : 1;
    const int per_spin = n_rho_cols / nspin;
    for (int sp = 0; per_spin > 0 && sp < nspin; ++sp) {
        std::vector<double> rho(4 * (size_t)n);
        if (int rc = download_rho(c, rho.data(), sp ? c->d_rho_b : c->d_rho)) return rc;
        for (int64_t p = 0; p < n; ++p)
            for (int q = 0; q < per_spin; ++q) m[(5 + per_spin * sp + q) * n + p] = rho[4 * p + q];
    }
    return pdfxc_save_general_matrix(path, n, nc, m.data());
}

}  // extern "C"
