// restart.cpp -- the reference's restart file of the flow field in the
// layout fsrc/restart.f03:33-98 reads ("old_res_form"): three list-directed
// header records (nscal; ti; bbx, dbbx), then per cell (i fastest over
// 0..jx+1, 0..jy+1, 0..jz+1) one record u v w fi rhov(1..nscal), then one
// record per contravariant flux uc (0..jx,1..jy,1..jz), vc (1..jx,0..jy,1..jz),
// wc (1..jx,1..jy,0..jz); data records in Fortran format (10e18.10): up to ten
// 18-character E fields per line.  Host code only; lets a served case start
// from a pressure field PLES wrote to disk.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ples_mg.h"

namespace {

// One value as Fortran's E18.10 edit descriptor writes it: 0.dddddddddd E+xx,
// right-justified in 18 columns; exponents beyond two digits drop the E.
void fortran_e18_10(double v, char out[19]) {
    char body[40];
    if (v == 0.0 || !std::isfinite(v)) {
        if (!std::isfinite(v)) std::snprintf(body, sizeof body, "%s", std::isnan(v) ? "NaN" : (v > 0 ? "Infinity" : "-Infinity"));
        else std::snprintf(body, sizeof body, "%s0.0000000000E+00", std::signbit(v) ? "-" : "");
    } else {
        char sci[40];
        std::snprintf(sci, sizeof sci, "%.9E", std::fabs(v));  // d.dddddddddE+xx, 10 significant digits
        const int e10 = std::atoi(std::strchr(sci, 'E') + 1) + 1;
        char mant[16];
        mant[0] = sci[0];
        std::memcpy(mant + 1, sci + 2, 9);
        mant[10] = 0;
        const char* sg = v < 0 ? "-" : "";
        if (std::abs(e10) <= 99)
            std::snprintf(body, sizeof body, "%s0.%sE%c%02d", sg, mant, e10 < 0 ? '-' : '+', std::abs(e10));
        else
            std::snprintf(body, sizeof body, "%s0.%s%c%03d", sg, mant, e10 < 0 ? '-' : '+', std::abs(e10));
    }
    std::snprintf(out, 19, "%18s", body);
}

struct RecWriter {
    FILE* f;
    std::vector<double> rec;
    void flush() {
        char fld[19];
        for (size_t q = 0; q < rec.size(); ++q) {
            fortran_e18_10(rec[q], fld);
            std::fputs(fld, f);
            if (q % 10 == 9 || q + 1 == rec.size()) std::fputc('\n', f);
        }
        rec.clear();
    }
};

// Reads 18-column E fields record by record (a record of n values spans
// ceil(n/10) lines).
struct RecReader {
    FILE* f;
    bool ok = true;
    bool field(const std::string& line, size_t q, double& v) {
        if (line.size() < 18 * (q + 1)) return false;
        std::string s = line.substr(18 * q, 18);
        std::string t;
        for (char c : s)
            if (!std::isspace((unsigned char)c)) t += (c == 'D' || c == 'd') ? 'E' : c;
        if (t.empty()) return false;
        // a three-digit exponent is written without its E: 0.1234567890+100
        if (t.find('E') == std::string::npos && t.find('e') == std::string::npos) {
            const size_t p = t.find_first_of("+-", 1);
            if (p != std::string::npos) t.insert(p, "E");
        }
        char* end = nullptr;
        v = std::strtod(t.c_str(), &end);
        return end && *end == 0;
    }
    bool record(size_t n, double* out) {
        char buf[512];
        size_t got = 0;
        while (got < n) {
            if (!std::fgets(buf, sizeof buf, f)) return ok = false;
            std::string line(buf);
            while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
            const size_t per = std::min<size_t>(10, n - got);
            for (size_t q = 0; q < per; ++q)
                if (!field(line, q, out[got + q])) return ok = false;
            got += per;
        }
        return true;
    }
};

}  // namespace

extern "C" {

int ples_mg_restart_write(const char* path, int jx, int jy, int jz, int nscal, double ti, double bbx, double dbbx,
                          const double* u, const double* v, const double* w, const double* fi, const double* rhov,
                          const double* uc, const double* vc, const double* wc) {
    if (!path || jx < 1 || jy < 1 || jz < 1 || nscal < 0 || !u || !v || !w || !fi || (nscal > 0 && !rhov) || !uc ||
        !vc || !wc)
        return PLES_MG_ERR_ARG;
    FILE* f = std::fopen(path, "w");
    if (!f) return PLES_MG_ERR_ARG;
    std::fprintf(f, " %d\n %.17G\n %.17G %.17G\n", nscal, ti, bbx, dbbx);
    RecWriter w_{f, {}};
    const long sx = jx + 2, sy = (long)(jx + 2) * (jy + 2);
    for (int k = 0; k <= jz + 1; ++k)
        for (int j = 0; j <= jy + 1; ++j)
            for (int i = 0; i <= jx + 1; ++i) {
                const long c = i + sx * j + sy * k;
                w_.rec = {u[c], v[c], w[c], fi[c]};
                for (int s = 0; s < nscal; ++s) w_.rec.push_back(rhov[s + (long)nscal * c]);
                w_.flush();
            }
    for (int k = 1; k <= jz; ++k)  // uc (0:jx, 1:jy, 1:jz)
        for (int j = 1; j <= jy; ++j)
            for (int i = 0; i <= jx; ++i) {
                w_.rec = {uc[i + (long)(jx + 1) * ((j - 1) + (long)jy * (k - 1))]};
                w_.flush();
            }
    for (int k = 1; k <= jz; ++k)  // vc (1:jx, 0:jy, 1:jz)
        for (int j = 0; j <= jy; ++j)
            for (int i = 1; i <= jx; ++i) {
                w_.rec = {vc[(i - 1) + (long)jx * (j + (long)(jy + 1) * (k - 1))]};
                w_.flush();
            }
    for (int k = 0; k <= jz; ++k)  // wc (1:jx, 1:jy, 0:jz)
        for (int j = 1; j <= jy; ++j)
            for (int i = 1; i <= jx; ++i) {
                w_.rec = {wc[(i - 1) + (long)jx * ((j - 1) + (long)jy * k)]};
                w_.flush();
            }
    const bool bad = std::ferror(f) != 0;
    std::fclose(f);
    return bad ? PLES_MG_ERR_ARG : PLES_MG_OK;
}

int ples_mg_restart_read(const char* path, int jx, int jy, int jz, int nscal, double* ti, double* bbx, double* dbbx,
                         double* u, double* v, double* w, double* fi, double* rhov, double* uc, double* vc,
                         double* wc) {
    if (!path || jx < 1 || jy < 1 || jz < 1 || nscal < 0) return PLES_MG_ERR_ARG;
    FILE* f = std::fopen(path, "r");
    if (!f) return PLES_MG_ERR_ARG;
    int ns = -1;
    double t = 0, b1 = 0, b2 = 0;
    if (std::fscanf(f, "%d %lf %lf %lf", &ns, &t, &b1, &b2) != 4 || ns != nscal) {
        std::fclose(f);
        return PLES_MG_ERR_ARG;
    }
    int ch;
    while ((ch = std::fgetc(f)) != EOF && ch != '\n') {
    }
    if (ti) *ti = t;
    if (bbx) *bbx = b1;
    if (dbbx) *dbbx = b2;
    RecReader r{f};
    std::vector<double> rec(4 + nscal);
    const long sx = jx + 2, sy = (long)(jx + 2) * (jy + 2);
    for (int k = 0; k <= jz + 1 && r.ok; ++k)
        for (int j = 0; j <= jy + 1 && r.ok; ++j)
            for (int i = 0; i <= jx + 1 && r.ok; ++i) {
                if (!r.record(rec.size(), rec.data())) break;
                const long c = i + sx * j + sy * k;
                if (u) u[c] = rec[0];
                if (v) v[c] = rec[1];
                if (w) w[c] = rec[2];
                if (fi) fi[c] = rec[3];
                if (rhov)
                    for (int s = 0; s < nscal; ++s) rhov[s + (long)nscal * c] = rec[4 + s];
            }
    double x = 0;
    for (int k = 1; k <= jz && r.ok; ++k)
        for (int j = 1; j <= jy && r.ok; ++j)
            for (int i = 0; i <= jx && r.ok; ++i)
                if (r.record(1, &x) && uc) uc[i + (long)(jx + 1) * ((j - 1) + (long)jy * (k - 1))] = x;
    for (int k = 1; k <= jz && r.ok; ++k)
        for (int j = 0; j <= jy && r.ok; ++j)
            for (int i = 1; i <= jx && r.ok; ++i)
                if (r.record(1, &x) && vc) vc[(i - 1) + (long)jx * (j + (long)(jy + 1) * (k - 1))] = x;
    for (int k = 0; k <= jz && r.ok; ++k)
        for (int j = 1; j <= jy && r.ok; ++j)
            for (int i = 1; i <= jx && r.ok; ++i)
                if (r.record(1, &x) && wc) wc[(i - 1) + (long)jx * ((j - 1) + (long)jy * k)] = x;
    std::fclose(f);
    return r.ok ? PLES_MG_OK : PLES_MG_ERR_ARG;
}

}  // extern "C"
