// io.cpp -- LAMMPS data files in and LAMMPS dump frames out (host code).
//
// Reference behaviour followed (paths relative to the reference checkout):
//   data file header and sections   rbmd-gpu-clone/rbmd-gpu/src/reader/src/atomic_reader.cpp:24-58
//   Atoms lines by atom style       atomic_reader.cpp:135-230 (atomic: id type x y z;
//                                   charge: id type q x y z; full: id mol type q x y z);
//                                   the atom with id k goes to index k - 1, types 0-based
//   Velocities lines                atomic_reader.cpp:322-350 (id vx vy vz)
//   trajectory frames               rbmd-parallel/framework/output/TrajectoryOutput.cpp:21-44
//                                   (ITEM: TIMESTEP / NUMBER OF ATOMS / BOX BOUNDS pp pp pp /
//                                   ATOMS id type x y z, ids and types 1-based)
// Beyond the reference: trailing image flags of Atoms lines are kept, every id
// must appear once (the reference writes out of range otherwise), and Masses and
// Pair Coeffs (lj/cut/coul/long: type eps sigma) are returned per type.
#include <stdint.h>
#include <stdio.h>
#include <ctype.h>
#include <errno.h>
#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <sstream>
#include <climits>
#include <string>
#include <vector>

#include "../../include/rbmd_force.h"

namespace {

thread_local std::string g_err;

int io_fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

// Text before a '#' comment, without surrounding blanks.
std::string strip(const std::string& s)
{
    std::string t = s.substr(0, s.find('#'));
    const size_t a = t.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return std::string();
    const size_t b = t.find_last_not_of(" \t\r\n");
    return t.substr(a, b - a + 1);
}

bool starts_with(const std::string& s, const char* p) { return s.compare(0, strlen(p), p) == 0; }

// Number tokens of a line, parsed in place (strtoll / strtod: correctly rounded);
// false when the next token is missing or not a whole number of that kind.
struct Tok {
    const char* p;
    bool i64(int64_t& v)
    {
        char* e;
        errno = 0;
        const long long t = strtoll(p, &e, 10);
        if (e == p || errno || !(*e == 0 || isspace((unsigned char)*e))) return false;
        v = t;
        p = e;
        return true;
    }
    bool i32(int& v)
    {
        int64_t t;
        if (!i64(t) || t < INT32_MIN || t > INT32_MAX) return false;
        v = (int)t;
        return true;
    }
    bool f64(double& v)
    {
        char* e;
        const double t = strtod(p, &e);
        if (e == p || !(*e == 0 || isspace((unsigned char)*e))) return false;
        v = t;
        p = e;
        return true;
    }
};

struct Header {
    int64_t natoms = -1;
    int32_t ntypes = -1;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    bool box[3] = {false, false, false};
};

// Header lines before the first section: "N atoms", "M atom types", "lo hi xlo xhi".
// Returns the stream positioned at the first section name (in *section).
int read_header(std::ifstream& in, Header& h, std::string* section, int64_t* lineno)
{
    std::string raw;
    bool first = true;
    while (std::getline(in, raw)) {
        ++*lineno;
        if (first) {  // the first line is a title
            first = false;
            continue;
        }
        const std::string s = strip(raw);
        if (s.empty()) continue;
        std::istringstream is(s);
        std::vector<std::string> tok;
        for (std::string t; is >> t;) tok.push_back(t);
        if (tok.size() == 2 && tok[1] == "atoms") {
            h.natoms = std::stoll(tok[0]);
        } else if (tok.size() == 3 && tok[1] == "atom" && tok[2] == "types") {
            h.ntypes = std::stoi(tok[0]);
        } else if (tok.size() == 4 && tok[2].size() == 3 && tok[3].size() == 3 &&
                   tok[2].substr(1) == "lo" && tok[3].substr(1) == "hi" && tok[2][0] == tok[3][0] &&
                   (tok[2][0] == 'x' || tok[2][0] == 'y' || tok[2][0] == 'z')) {
            const int d = tok[2][0] - 'x';
            h.lo[d] = std::stod(tok[0]);
            h.hi[d] = std::stod(tok[1]);
            h.box[d] = true;
        } else if (isalpha((unsigned char)s[0])) {
            *section = s;
            return RBMD_OK;
        }
        // other header keywords (bonds, angles, tilt factors ...) are not used
    }
    *section = std::string();
    return RBMD_OK;
}

int columns_of(const char* style)
{
    if (!style || !strcmp(style, "charge")) return 6;
    if (!strcmp(style, "atomic")) return 5;
    if (!strcmp(style, "full")) return 7;
    return -1;
}

}  // namespace

extern "C" {

const char* rbmd_io_last_error(void) { return g_err.c_str(); }

int rbmd_read_data_info(const char* path, rbmd_data_info* info)
{
    if (!path || !info) return io_fail(RBMD_ERR_INVALID, "read_data: null argument");
    std::ifstream in(path);
    if (!in) return io_fail(RBMD_ERR_INVALID, std::string("read_data: cannot open ") + path);
    Header h;
    std::string sec;
    int64_t lineno = 0;
    try {
        read_header(in, h, &sec, &lineno);
    } catch (const std::exception&) {
        return io_fail(RBMD_ERR_INVALID, "read_data: malformed header line " +
                                             std::to_string(lineno));
    }
    if (h.natoms < 0 || h.ntypes < 1 || !h.box[0] || !h.box[1] || !h.box[2])
        return io_fail(RBMD_ERR_INVALID,
                       "read_data: header needs 'N atoms', 'M atom types' and xlo/ylo/zlo lines");
    memset(info, 0, sizeof *info);
    info->natoms = h.natoms;
    info->ntypes = h.ntypes;
    for (int d = 0; d < 3; ++d) {
        info->lo[d] = h.lo[d];
        info->hi[d] = h.hi[d];
    }
    for (std::string raw; !sec.empty() || std::getline(in, raw);) {
        const std::string s = sec.empty() ? strip(raw) : sec;
        sec.clear();
        if (starts_with(s, "Masses")) info->has_masses = 1;
        else if (starts_with(s, "Pair Coeffs")) info->has_pair_coeffs = 1;
        else if (starts_with(s, "Velocities")) info->has_velocities = 1;
    }
    return RBMD_OK;
}

int rbmd_read_data(const char* path, const char* atom_style, int64_t n, int32_t* type,
                   double* q, double* x, double* y, double* z, int32_t* img, double* vx,
                   double* vy, double* vz, double* mass, double* eps, double* sigma)
{
    const int cols = columns_of(atom_style);
    if (cols < 0) return io_fail(RBMD_ERR_INVALID, "read_data: atom_style must be atomic, charge or full");
    if (!path || !type || !x || !y || !z) return io_fail(RBMD_ERR_INVALID, "read_data: null argument");
    std::ifstream in(path);
    if (!in) return io_fail(RBMD_ERR_INVALID, std::string("read_data: cannot open ") + path);
    Header h;
    std::string sec;
    int64_t lineno = 0;
    try {
        read_header(in, h, &sec, &lineno);
    } catch (const std::exception&) {
        return io_fail(RBMD_ERR_INVALID, "read_data: malformed header line " + std::to_string(lineno));
    }
    if (h.natoms != n) return io_fail(RBMD_ERR_INVALID, "read_data: n differs from the file's atom count");
    std::vector<char> seen((size_t)n, 0);
    int64_t natoms_read = 0;
    enum { S_NONE, S_MASS, S_PAIR, S_ATOMS, S_VEL, S_OTHER } mode = S_NONE;
    std::string raw;
    while (true) {
        std::string s;
        if (!sec.empty()) {
            s = sec;
            sec.clear();
        } else {
            if (!std::getline(in, raw)) break;
            ++lineno;
            s = strip(raw);
        }
        if (s.empty()) continue;
        if (isalpha((unsigned char)s[0])) {
            if (starts_with(s, "Masses")) mode = S_MASS;
            else if (starts_with(s, "Pair Coeffs")) mode = S_PAIR;
            else if (starts_with(s, "Atoms")) mode = S_ATOMS;
            else if (starts_with(s, "Velocities")) mode = S_VEL;
            else mode = S_OTHER;  // Bonds, Angles, ... are not part of this path
            continue;
        }
        Tok is{s.c_str()};
        auto where = [&]() { return "read_data: line " + std::to_string(lineno) + ": "; };
        if (mode == S_MASS || mode == S_PAIR) {
            int t;
            double a, b = 0;
            if (!is.i32(t) || !is.f64(a) || (mode == S_PAIR && !is.f64(b)) || t < 1 ||
                t > h.ntypes)
                return io_fail(RBMD_ERR_INVALID, where() + "bad type line");
            if (mode == S_MASS && mass) mass[t - 1] = a;
            if (mode == S_PAIR) {
                if (eps) eps[t - 1] = a;
                if (sigma) sigma[t - 1] = b;
            }
        } else if (mode == S_ATOMS) {
            int64_t id, mol;
            int t;
            double qq = 0, p[3];
            if (!is.i64(id) || (cols == 7 && !is.i64(mol)) || !is.i32(t) ||
                (cols != 5 && !is.f64(qq)) || !is.f64(p[0]) || !is.f64(p[1]) || !is.f64(p[2]))
                return io_fail(RBMD_ERR_INVALID, where() + "bad atom line");
            if (id < 1 || id > n) return io_fail(RBMD_ERR_INVALID, where() + "atom id out of range");
            if (t < 1 || t > h.ntypes)
                return io_fail(RBMD_ERR_INVALID, where() + "atom type out of range");
            const int64_t i = id - 1;
            if (seen[i]) return io_fail(RBMD_ERR_INVALID, where() + "atom id repeated");
            seen[i] = 1;
            ++natoms_read;
            type[i] = t - 1;
            if (q) q[i] = qq;
            x[i] = p[0];
            y[i] = p[1];
            z[i] = p[2];
            int im[3] = {0, 0, 0};
            if (is.i32(im[0]) && !(is.i32(im[1]) && is.i32(im[2])))
                return io_fail(RBMD_ERR_INVALID, where() + "bad image flags");
            if (img) {
                img[3 * i] = im[0];
                img[3 * i + 1] = im[1];
                img[3 * i + 2] = im[2];
            }
        } else if (mode == S_VEL) {
            int64_t id;
            double v[3];
            if (!is.i64(id) || !is.f64(v[0]) || !is.f64(v[1]) || !is.f64(v[2]))
                return io_fail(RBMD_ERR_INVALID, where() + "bad velocity line");
            if (id < 1 || id > n)
                return io_fail(RBMD_ERR_INVALID, where() + "velocity id out of range");
            if (vx) vx[id - 1] = v[0];
            if (vy) vy[id - 1] = v[1];
            if (vz) vz[id - 1] = v[2];
        }
    }
    if (natoms_read != n) return io_fail(RBMD_ERR_INVALID, "read_data: Atoms section lacks some ids");
    return RBMD_OK;
}

int rbmd_write_dump(const char* path, int32_t append, int64_t step, int64_t n, const double* lo,
                    const double* hi, const int32_t* type, const double* x, const double* y,
                    const double* z)
{
    if (!path || !lo || !hi || (n > 0 && (!type || !x || !y || !z)))
        return io_fail(RBMD_ERR_INVALID, "write_dump: null argument");
    FILE* f = fopen(path, append ? "a" : "w");
    if (!f) return io_fail(RBMD_ERR_INVALID, std::string("write_dump: cannot open ") + path);
    fprintf(f, "ITEM: TIMESTEP\n%lld\nITEM: NUMBER OF ATOMS\n%lld\nITEM: BOX BOUNDS pp pp pp\n",
            (long long)step, (long long)n);
    for (int d = 0; d < 3; ++d) fprintf(f, "%.17g %.17g\n", lo[d], hi[d]);
    fprintf(f, "ITEM: ATOMS id type x y z\n");
    for (int64_t i = 0; i < n; ++i)
        fprintf(f, "%lld %d %.17g %.17g %.17g\n", (long long)(i + 1), type[i] + 1, x[i], y[i], z[i]);
    const int bad = ferror(f);
    if (fclose(f) != 0 || bad) return io_fail(RBMD_ERR_INVALID, "write_dump: write failed");
    return RBMD_OK;
}

}  // extern "C"
