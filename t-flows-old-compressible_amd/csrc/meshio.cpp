// meshio.cpp -- the T-FlowS mesh files on the host, and the system pattern.
//
//   NAME.cns  connectivity, read as Process/CnsLoa.f90:38-62 reads it
//             (written by Library/GeoSav.f90:16-112): NC, NbC, NS, NSsh,
//             Nmat; material(1:NC), material(-1:-NbC:-1); SideC(0,:),
//             SideC(1,:), SideC(2,:); bcmark(-1:-NbC:-1); CopyC(-1:-NbC:-1)
//   NAME.geo  geometry, read as Library/GeoLoa.f90:22-49: xc, yc, zc of the
//             cells 1..NC and of the boundary cells -1..-NbC, volume, delta,
//             WallDs (1..NC), Sx, Sy, Sz, Dx, Dy, Dz, f, xsp, ysp, zsp (NS)
//
// Both are Fortran sequential unformatted files: each record is framed by
// its byte count as a 4-byte integer before and after it; a record longer
// than 2^31 - 1 bytes is split into subrecords whose leading count is
// negated on all but the last.  INTEGER is 4 bytes and REAL 8 bytes, the
// kinds of the reference build (-fdefault-real-8); a record of 4-byte reals
// or 8-byte integers (another build's kinds) is widened or narrowed by its
// byte count.
//
// From the connectivity, the pattern of the system matrix as
// Process/TopolM.f90 builds it: per row the diagonal and one entry per
// interior face, columns sorted ascending, Adia the diagonal's position,
// SidAij(1,s) / SidAij(2,s) the positions of A(c1,c2) / A(c2,c1) -- where
// two faces join the same two cells, both take the last matching position
// of the row, as TopolM's search does (:65-74).  And from the geometry,
// what Process/Calc3.f90:9-49 derives for the assemblies: the centre
// distance D = x(c2) + D_file - x(c1) (D_file carries the periodic shift),
// Scoef = |S|^2 / (S . D), and fF (f; 1 or 0 on fluid-solid faces; 1 on
// non-BUFFER boundary faces).  TypeBC comes from bcmark and the
// boundary-type table of the .b file as Process/BouLoa.f90:167-175, 292-294
// sets it.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <cctype>
#include <cstdlib>
#include <vector>

#include "../../include/tflows_krylov.h"

namespace {

constexpr int kBuffer = 11, kFluid = 7, kSolid = 8;   // Library/allp_mod.f90:23-28

class RecordFile {
public:
    explicit RecordFile(const char* path) : path_(path ? path : "") {
        if (!path) throw std::invalid_argument("null file name");
        f_ = std::fopen(path, "rb");
        if (!f_) throw std::runtime_error(std::string("cannot open ") + path);
        // the file's size bounds every record length before any buffer grows
        if (std::fseek(f_, 0, SEEK_END) != 0) fail("cannot seek");
        const long sz = std::ftell(f_);
        if (sz < 0 || std::fseek(f_, 0, SEEK_SET) != 0) fail("cannot seek");
        size_ = (uint64_t)sz;
    }
    ~RecordFile() {
        if (f_) std::fclose(f_);
    }
    RecordFile(const RecordFile&) = delete;
    RecordFile& operator=(const RecordFile&) = delete;

    // The next record's bytes (subrecords joined).
    std::vector<unsigned char> next() {
        std::vector<unsigned char> out;
        for (;;) {
            int32_t head = 0, tail = 0;
            if (std::fread(&head, 4, 1, f_) != 1) fail("end of file where a record was expected");
            const bool more = head < 0;
            const uint64_t len = more ? (uint64_t)(-(int64_t)head) : (uint64_t)head;
            const long pos = std::ftell(f_);
            if (pos < 0 || len + 4 > size_ - (uint64_t)pos) fail("record shorter than its count");
            const size_t at = out.size();
            out.resize(at + len);
            if (len && std::fread(out.data() + at, 1, len, f_) != len) fail("record shorter than its count");
            if (std::fread(&tail, 4, 1, f_) != 1) fail("record without a closing count");
            if ((uint64_t)(tail < 0 ? -(int64_t)tail : tail) != len) fail("record counts disagree");
            if (!more) return out;
        }
    }

    std::vector<int32_t> ints(size_t count, const char* what) {
        const std::vector<unsigned char> r = next();
        std::vector<int32_t> v(count);
        if (r.size() == 4 * count) {
            if (count) std::memcpy(v.data(), r.data(), r.size());
        } else if (r.size() == 8 * count) {
            for (size_t i = 0; i < count; ++i) {
                int64_t x;
                std::memcpy(&x, r.data() + 8 * i, 8);
                if (x < INT32_MIN || x > INT32_MAX) fail(std::string(what) + ": integer out of range");
                v[i] = (int32_t)x;
            }
        } else {
            fail(std::string(what) + ": record of " + std::to_string(r.size()) + " bytes for " +
                 std::to_string(count) + " integers");
        }
        return v;
    }

    int32_t one_int(const char* what) { return ints(1, what)[0]; }

    std::vector<double> reals(size_t count, const char* what) {
        const std::vector<unsigned char> r = next();
        std::vector<double> v(count);
        if (r.size() == 8 * count) {
            if (count) std::memcpy(v.data(), r.data(), r.size());
        } else if (r.size() == 4 * count) {
            for (size_t i = 0; i < count; ++i) {
                float x;
                std::memcpy(&x, r.data() + 4 * i, 4);
                v[i] = x;
            }
        } else {
            fail(std::string(what) + ": record of " + std::to_string(r.size()) + " bytes for " +
                 std::to_string(count) + " reals");
        }
        return v;
    }

private:
    [[noreturn]] void fail(const std::string& msg) { throw std::runtime_error(path_ + ": " + msg); }
    std::string path_;
    uint64_t size_ = 0;
    std::FILE* f_ = nullptr;
};

}  // namespace

struct tflows_mesh_s {
    int32_t nc = 0, nbc = 0, ns = 0, nsh = 0, nmat = 0, nonz = 0;
    bool has_geo = false;
    // (-NbC:NC) arrays at [NbC + c]
    std::vector<int32_t> material, bcmark_cells, copyc_cells;
    std::vector<int32_t> sidec;          // SideC(0:2,NS), column-major: [3 (s-1) + k]
    std::vector<double> xc, yc, zc;      // (-NbC:NC)
    std::vector<double> volume, delta, wallds;   // (-NbC:NC), cells 1..NC from the file
    std::vector<double> sx, sy, sz, dxf, dyf, dzf, f, xsp, ysp, zsp;   // (NS), as in the file
    std::vector<double> dx, dy, dz, scoef;                             // Calc3
    // TopolM
    std::vector<int32_t> acol, arow, adia, sidaij;   // 1-based; SidAij(1:2,NS) column-major
};

namespace {

void load_cns(tflows_mesh_s& m, const char* path) {
    RecordFile r(path);
    m.nc = r.one_int("NC");
    m.nbc = r.one_int("NbC");
    m.ns = r.one_int("NS");
    m.nsh = r.one_int("NSsh");
    m.nmat = r.one_int("Nmat");
    if (m.nc <= 0 || m.nbc < 0 || m.ns < 0) throw std::runtime_error(std::string(path) + ": bad sizes");
    const size_t nall = (size_t)m.nbc + m.nc + 1;
    m.material.assign(nall, 0);
    const std::vector<int32_t> mi = r.ints(m.nc, "material(1:NC)");
    for (int c = 1; c <= m.nc; ++c) m.material[m.nbc + c] = mi[c - 1];
    const std::vector<int32_t> mb = r.ints(m.nbc, "material(-1:-NbC)");
    for (int k = 0; k < m.nbc; ++k) m.material[m.nbc - 1 - k] = mb[k];        // c = -1 - k
    m.sidec.assign(3 * (size_t)m.ns, 0);
    for (int k = 0; k < 3; ++k) {
        const std::vector<int32_t> v = r.ints(m.ns, "SideC");
        for (int s = 0; s < m.ns; ++s) m.sidec[3 * (size_t)s + k] = v[s];
    }
    m.bcmark_cells.assign(nall, 0);
    m.copyc_cells.assign(nall, 0);
    const std::vector<int32_t> bm = r.ints(m.nbc, "bcmark");
    const std::vector<int32_t> cc = r.ints(m.nbc, "CopyC");
    for (int k = 0; k < m.nbc; ++k) {
        m.bcmark_cells[m.nbc - 1 - k] = bm[k];
        m.copyc_cells[m.nbc - 1 - k] = cc[k];
    }
    for (int s = 0; s < m.ns; ++s) {
        const int c1 = m.sidec[3 * (size_t)s + 1], c2 = m.sidec[3 * (size_t)s + 2];
        if (c1 < 1 || c1 > m.nc || c2 < -m.nbc || c2 > m.nc || c2 == 0)
            throw std::runtime_error(std::string(path) + ": SideC out of range at face " + std::to_string(s + 1));
    }
}

void load_geo(tflows_mesh_s& m, const char* path) {
    RecordFile r(path);
    const size_t nall = (size_t)m.nbc + m.nc + 1;
    std::vector<double>* xyz[3] = {&m.xc, &m.yc, &m.zc};
    for (auto* a : xyz) a->assign(nall, 0.0);
    for (int k = 0; k < 3; ++k) {
        const std::vector<double> v = r.reals(m.nc, "xc(1:NC)");
        for (int c = 1; c <= m.nc; ++c) (*xyz[k])[m.nbc + c] = v[c - 1];
    }
    for (int k = 0; k < 3; ++k) {
        const std::vector<double> v = r.reals(m.nbc, "xc(-1:-NbC)");
        for (int q = 0; q < m.nbc; ++q) (*xyz[k])[m.nbc - 1 - q] = v[q];
    }
    std::vector<double>* cell[3] = {&m.volume, &m.delta, &m.wallds};
    for (auto* a : cell) {
        a->assign(nall, 0.0);
        const std::vector<double> v = r.reals(m.nc, "volume/delta/WallDs");
        for (int c = 1; c <= m.nc; ++c) (*a)[m.nbc + c] = v[c - 1];
    }
    std::vector<double>* face[10] = {&m.sx, &m.sy, &m.sz, &m.dxf, &m.dyf, &m.dzf, &m.f, &m.xsp, &m.ysp, &m.zsp};
    for (auto* a : face) *a = r.reals(m.ns, "face array");
    m.has_geo = true;
}

// The matrix pattern from the face list, built as a sort of (row, column)
// keys: one key per cell for its diagonal and two per interior face,
// (c1, c2) and (c2, c1), ordered by row, then column.  Row c then holds the
// diagonal and one entry per face of c in ascending column order, the
// layout of Process/TopolM.f90 (its ISORT, :72-74), with NONZERO = keys + 1
// (:49).  A key's position is where TopolM's searches find it (:75-94): the
// last entry of its row that holds the column, so two faces that join the
// same two cells both receive the later of their positions and the
// assemblies accumulate there.
void topolm(tflows_mesh_s& m) {
    const int nc = m.nc, ns = m.ns;
    std::vector<int> inner;                          // 0-based interior faces
    for (int s = 0; s < ns; ++s)
        if (m.sidec[3 * (size_t)s + 2] > 0) inner.push_back(s);
    const size_t nkeys = (size_t)nc + 2 * inner.size();
    if (nkeys + 1 > (size_t)INT32_MAX) throw std::runtime_error("NONZERO exceeds 32 bits");
    // key = row << 32 | column (1-based cells); origin k < nc: diagonal of
    // cell k + 1; nc + 2 f (+1): face inner[f], side 1 (side 2)
    std::vector<uint64_t> key(nkeys);
    for (int c = 1; c <= nc; ++c) key[c - 1] = (uint64_t)c << 32 | (uint32_t)c;
    for (size_t f = 0; f < inner.size(); ++f) {
        const uint64_t c1 = (uint32_t)m.sidec[3 * (size_t)inner[f] + 1];
        const uint64_t c2 = (uint32_t)m.sidec[3 * (size_t)inner[f] + 2];
        key[nc + 2 * f] = c1 << 32 | c2;
        key[nc + 2 * f + 1] = c2 << 32 | c1;
    }
    std::vector<uint32_t> order(nkeys);
    for (size_t k = 0; k < nkeys; ++k) order[k] = (uint32_t)k;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
    // where each origin lands, taken to the last position of its key's run
    std::vector<int32_t> at(nkeys);
    size_t run_end = nkeys;
    for (size_t p = nkeys; p-- > 0;) {
        if (p + 1 == nkeys || key[order[p]] != key[order[p + 1]]) run_end = p;
        at[order[p]] = (int32_t)run_end + 1;          // 1-based
    }
    m.nonz = (int32_t)nkeys + 1;
    m.arow.assign(nkeys + 1, 0);
    m.acol.assign((size_t)nc + 1, 0);
    m.adia.assign(nc, 0);
    for (size_t p = 0; p < nkeys; ++p) {
        m.arow[p] = (int32_t)(key[order[p]] & 0xffffffffu);
        m.acol[key[order[p]] >> 32]++;                // entries of each row
    }
    m.acol[0] = 1;
    for (int c = 1; c <= nc; ++c) m.acol[c] += m.acol[c - 1];
    for (int c = 1; c <= nc; ++c) m.adia[c - 1] = at[c - 1];
    m.sidaij.assign(2 * (size_t)ns, 0);
    for (size_t f = 0; f < inner.size(); ++f) {
        m.sidaij[2 * (size_t)inner[f]] = at[nc + 2 * f];
        m.sidaij[2 * (size_t)inner[f] + 1] = at[nc + 2 * f + 1];
    }
}

// Process/Calc3.f90:9-30: the centre distance and Scoef
void calc3(tflows_mesh_s& m) {
    const int ns = m.ns, nb = m.nbc;
    m.dx.resize(ns);
    m.dy.resize(ns);
    m.dz.resize(ns);
    m.scoef.resize(ns);
    for (int s = 0; s < ns; ++s) {
        const int c1 = nb + m.sidec[3 * (size_t)s + 1], c2 = nb + m.sidec[3 * (size_t)s + 2];
        const double sc = (m.sx[s] * m.sx[s] + m.sy[s] * m.sy[s]) + m.sz[s] * m.sz[s];
        const double xc2 = m.xc[c2] + m.dxf[s], yc2 = m.yc[c2] + m.dyf[s], zc2 = m.zc[c2] + m.dzf[s];
        m.dx[s] = xc2 - m.xc[c1];
        m.dy[s] = yc2 - m.yc[c1];
        m.dz[s] = zc2 - m.zc[c1];
        m.scoef[s] = sc / ((m.dx[s] * m.sx[s] + m.dy[s] * m.sy[s]) + m.dz[s] * m.sz[s]);
    }
}

template <class T>
int copy_out(const std::vector<T>& v, T* dst, int64_t count) {
    if (!dst) throw std::invalid_argument("null destination");
    if (count != (int64_t)v.size())
        throw std::invalid_argument("array holds " + std::to_string(v.size()) + " entries, not " +
                                    std::to_string(count));
    if (count) std::memcpy(dst, v.data(), sizeof(T) * (size_t)count);
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// NAME.b, the boundary-condition file (Process/BouLoa.f90:30-160): materials
// (state, VISc, DENc, and CONc, CAPc when HOT), boundary conditions in file
// order (type, then U V W [P for PRESSURE] [Zmix when HOT], or FILE and a
// profile name) and initial conditions (U V W [Zmix], or FILE and a name).
// Lines are read as Library/ReadC.f90 does: a line is read as A300, so
// columns past 300 are never seen; a line of spaces only is skipped
// (`string == ''`, :22), a line whose first column is one of
// ! # $ % / & [ ] { } : ; @ < > * is a comment, and tokens are separated by
// characters below '!'.  A line holding only tabs or other control
// characters is a data line without tokens, which the caller's next read
// refuses.  ReadC closes a token only when a separator follows it within
// column 299 (:44-51), so a token reaching column 299 is refused.  A
// carriage return ending the line (a CRLF file) is dropped.
// ---------------------------------------------------------------------------
struct tflows_bfile_s {
    int nmat = 0, nb = 0, ninit = 0;
    double visc = 0.0;                           // VISc: the last material's (:51)
    std::vector<int32_t> statemat, typbou, bfile, ifile;
    std::vector<double> denc, conc, capc, bu, bv, bw, bp, bz, iu, iv, iw, iz;
    std::vector<std::string> profile, initname;
};

namespace {

struct BLines {
    std::ifstream in;
    std::string path;
    int lineno = 0;
    std::vector<std::string> next(const char* what) {        // ReadC
        std::string line;
        while (std::getline(in, line)) {
            ++lineno;
            if (!line.empty() && line.back() == '\r') line.pop_back();
            if (line.size() > 300) line.resize(300);
            bool blank = true;
            for (char ch : line) blank = blank && ch == ' ';
            if (blank) continue;
            if (std::strchr("!#$%/&[]{}:;@<>*", line[0])) continue;
            if (line.size() > 298 && (unsigned char)line[298] >= '!')
                throw std::runtime_error(where() + ": a token reaches column 299 (ReadC reads 300 columns and "
                                                   "closes tokens only up to column 298)");
            std::vector<std::string> tok;
            size_t i = 0;
            while (i < line.size()) {
                while (i < line.size() && (unsigned char)line[i] < '!') ++i;
                size_t j = i;
                while (j < line.size() && (unsigned char)line[j] >= '!') ++j;
                if (j > i) tok.push_back(line.substr(i, j - i));
                i = j;
            }
            return tok;
        }
        throw std::runtime_error(path + ": end of file before " + what);
    }
    std::string where() const { return path + " line " + std::to_string(lineno); }
};

std::string upper(std::string t) {                          // ToUppr
    for (char& ch : t) ch = (char)std::toupper((unsigned char)ch);
    return t;
}

// a number as Fortran's list-directed read takes it (D exponents included)
double fnum(const BLines& L, const std::vector<std::string>& t, size_t k, const char* what) {
    if (k >= t.size()) throw std::runtime_error(L.where() + ": missing " + what);
    std::string v = t[k];
    for (char& ch : v)
        if (ch == 'd' || ch == 'D') ch = 'e';
    char* end = nullptr;
    const double x = std::strtod(v.c_str(), &end);
    if (end == v.c_str() || *end != '\0') throw std::runtime_error(L.where() + ": bad " + what + " '" + t[k] + "'");
    return x;
}

int fint(const BLines& L, const std::vector<std::string>& t, size_t k, const char* what) {
    const double x = fnum(L, t, k, what);
    if (x != (double)(int)x) throw std::runtime_error(L.where() + ": " + what + " is not an integer");
    return (int)x;
}

void load_b(tflows_bfile_s& B, const char* path, bool hot) {
    BLines L;
    L.path = path ? path : "";
    L.in.open(L.path);
    if (!L.in) throw std::runtime_error("cannot open " + L.path);
    auto t = L.next("the number of materials");
    B.nmat = fint(L, t, 0, "Nmat");                             // :40
    if (B.nmat < 1 || B.nmat > 1024) throw std::runtime_error(L.where() + ": Nmat out of range");
    for (int n = 0; n < B.nmat; ++n) {                          // :41-56
        t = L.next("a material");
        if (t.size() < 2) throw std::runtime_error(L.where() + ": missing material state");
        const std::string st = upper(t[1]);
        if (st == "FLUID") B.statemat.push_back(7);
        else if (st == "SOLID") B.statemat.push_back(8);
        else throw std::runtime_error(L.where() + ": unknown material state " + t[1]);
        B.visc = fnum(L, t, 2, "VISc");
        B.denc.push_back(fnum(L, t, 3, "DENc"));
        B.conc.push_back(hot ? fnum(L, t, 4, "CONc") : 0.0);
        B.capc.push_back(hot ? fnum(L, t, 5, "CAPc") : 0.0);
    }
    t = L.next("the number of boundary conditions");
    B.nb = fint(L, t, 0, "NB");                                 // :61
    if (B.nb < 0 || B.nb > 128) throw std::runtime_error(L.where() + ": NB out of range (0..128)");
    static const std::pair<const char*, int> kTypes[] = {{"INFLOW", 1},  {"WALL", 2},     {"OUTFLOW", 3},
                                                         {"SYMMETRY", 4}, {"WALLFLUX", 6}, {"CONVECTIVE", 5},
                                                         {"PRESSURE", 12}};
    for (int n = 0; n < B.nb; ++n) {                            // :62-124
        t = L.next("a boundary condition");
        fint(L, t, 0, "the boundary index");
        if (t.size() < 3) throw std::runtime_error(L.where() + ": incomplete boundary condition");
        const std::string ty = upper(t[1]);
        int code = 0;
        for (const auto& kv : kTypes)
            if (ty == kv.first) code = kv.second;
        if (!code) throw std::runtime_error(L.where() + ": unknown boundary condition type " + t[1]);
        B.typbou.push_back(code);
        double u = 0.0, v = 0.0, w = 0.0, pb = 0.0, z = 0.0;
        std::string prof;
        const bool file = upper(t[2]) == "FILE";
        if (file) {
            if (t.size() < 4) throw std::runtime_error(L.where() + ": FILE without a name");
            prof = t[3];
        } else {
            u = fnum(L, t, 2, "U");
            v = fnum(L, t, 3, "V");
            w = fnum(L, t, 4, "W");
            size_t k = 5;
            if (code == 12) pb = fnum(L, t, k++, "P");
            if (hot) z = fnum(L, t, k, "Zmix");
        }
        B.bfile.push_back(file ? 1 : 0);
        B.profile.push_back(prof);
        B.bu.push_back(u);
        B.bv.push_back(v);
        B.bw.push_back(w);
        B.bp.push_back(pb);
        B.bz.push_back(z);
    }
    t = L.next("the number of initial conditions");
    B.ninit = fint(L, t, 0, "Ninit");                           // :130
    if (B.ninit < 0 || B.ninit > 128) throw std::runtime_error(L.where() + ": Ninit out of range (0..128)");
    for (int n = 0; n < B.ninit; ++n) {                         // :135-160
        t = L.next("an initial condition");
        if (t.size() < 2) throw std::runtime_error(L.where() + ": incomplete initial condition");
        const bool file = upper(t[1]) == "FILE";
        double u = 0.0, v = 0.0, w = 0.0, z = 0.0;
        std::string nm;
        if (file) {
            if (t.size() < 3) throw std::runtime_error(L.where() + ": FILE without a name");
            nm = t[2];
        } else {
            u = fnum(L, t, 1, "U");
            v = fnum(L, t, 2, "V");
            w = fnum(L, t, 3, "W");
            if (hot) z = fnum(L, t, 4, "Zmix");
        }
        B.ifile.push_back(file ? 1 : 0);
        B.initname.push_back(nm);
        B.iu.push_back(u);
        B.iv.push_back(v);
        B.iw.push_back(w);
        B.iz.push_back(z);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// C API (errors through tflows_last_error, defined in capi.hip)
// ---------------------------------------------------------------------------
namespace tflows {
int set_last_error(const std::string& msg);
}

template <class F>
static int mesh_guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        return tflows::set_last_error(e.what());
    }
}

extern "C" {

int tflows_mesh_load(const char* cns_path, const char* geo_path, tflows_mesh* out) {
    return mesh_guard([&] {
        if (!out) throw std::invalid_argument("out is NULL");
        auto m = std::make_unique<tflows_mesh_s>();
        load_cns(*m, cns_path);
        topolm(*m);
        if (geo_path) {
            load_geo(*m, geo_path);
            calc3(*m);
        }
        *out = m.release();
    });
}

int tflows_mesh_destroy(tflows_mesh m) {
    return mesh_guard([&] { delete m; });
}

int tflows_mesh_info(tflows_mesh m, int32_t* nc, int32_t* nbc, int32_t* ns, int32_t* nmat, int32_t* nonz,
                     int32_t* has_geo) {
    return mesh_guard([&] {
        if (!m) throw std::invalid_argument("null mesh");
        if (nc) *nc = m->nc;
        if (nbc) *nbc = m->nbc;
        if (ns) *ns = m->ns;
        if (nmat) *nmat = m->nmat;
        if (nonz) *nonz = m->nonz;
        if (has_geo) *has_geo = m->has_geo ? 1 : 0;
    });
}

int tflows_mesh_int_array(tflows_mesh m, const char* name, int32_t* dst, int64_t count) {
    return mesh_guard([&] {
        if (!m || !name) throw std::invalid_argument("null mesh or name");
        const std::string n(name);
        if (n == "material") copy_out(m->material, dst, count);
        else if (n == "bcmark") copy_out(m->bcmark_cells, dst, count);
        else if (n == "copyc") copy_out(m->copyc_cells, dst, count);
        else if (n == "sidec") copy_out(m->sidec, dst, count);
        else if (n == "acol") copy_out(m->acol, dst, count);
        else if (n == "arow") copy_out(m->arow, dst, count);
        else if (n == "adia") copy_out(m->adia, dst, count);
        else if (n == "sidaij") copy_out(m->sidaij, dst, count);
        else throw std::invalid_argument("unknown integer array " + n);
    });
}

int tflows_mesh_real_array(tflows_mesh m, const char* name, double* dst, int64_t count) {
    return mesh_guard([&] {
        if (!m || !name) throw std::invalid_argument("null mesh or name");
        if (!m->has_geo) throw std::invalid_argument("the mesh was loaded without its .geo file");
        const std::string n(name);
        const std::pair<const char*, const std::vector<double>*> table[] = {
            {"xc", &m->xc},         {"yc", &m->yc},         {"zc", &m->zc},         {"volume", &m->volume},
            {"delta", &m->delta},   {"wallds", &m->wallds}, {"sx", &m->sx},         {"sy", &m->sy},
            {"sz", &m->sz},         {"dx_file", &m->dxf},   {"dy_file", &m->dyf},   {"dz_file", &m->dzf},
            {"f", &m->f},           {"xsp", &m->xsp},       {"ysp", &m->ysp},       {"zsp", &m->zsp},
            {"dx", &m->dx},         {"dy", &m->dy},         {"dz", &m->dz},         {"scoef", &m->scoef}};
        for (const auto& t : table)
            if (n == t.first) {
                copy_out(*t.second, dst, count);
                return;
            }
        throw std::invalid_argument("unknown real array " + n);
    });
}

int tflows_mesh_typebc(tflows_mesh m, const int32_t* typbou, int32_t nb, int32_t* typebc, int64_t count) {
    return mesh_guard([&] {
        if (!m || !typebc) throw std::invalid_argument("null argument");
        if (nb < 0 || (nb > 0 && !typbou)) throw std::invalid_argument("bad boundary-type table");
        const int64_t nall = (int64_t)m->nbc + m->nc + 1;
        if (count != nall) throw std::invalid_argument("TypeBC holds NbC + NC + 1 entries");
        std::fill(typebc, typebc + nall, 0);
        for (int c = -1; c >= -m->nbc; --c) {                          // BouLoa.f90:167-175
            const int mark = m->bcmark_cells[m->nbc + c];
            if (mark >= 1 && mark <= nb) typebc[m->nbc + c] = typbou[mark - 1];
        }
        for (int c = -1; c >= -m->nbc; --c)                            // :292-294
            if (m->bcmark_cells[m->nbc + c] == kBuffer) typebc[m->nbc + c] = kBuffer;
    });
}

int tflows_mesh_ff(tflows_mesh m, const int32_t* typebc, const int32_t* statemat, double* ff, int64_t count) {
    return mesh_guard([&] {
        if (!m || !typebc || !ff) throw std::invalid_argument("null argument");
        if (!m->has_geo) throw std::invalid_argument("the mesh was loaded without its .geo file");
        if (count != m->ns) throw std::invalid_argument("fF holds NS entries");
        auto state = [&](int c) {
            const int mat = m->material[m->nbc + c];
            if (!statemat) return kFluid;
            if (mat < 1 || mat > m->nmat) throw std::invalid_argument("material out of range");
            return statemat[mat - 1];
        };
        for (int s = 0; s < m->ns; ++s) {                              // Calc3.f90:53-71
            const int c1 = m->sidec[3 * (size_t)s + 1], c2 = m->sidec[3 * (size_t)s + 2];
            ff[s] = m->f[s];
            const int s1 = state(c1), s2 = state(c2);
            if (s1 == kFluid && s2 == kSolid) ff[s] = 1.0;
            if (s1 == kSolid && s2 == kFluid) ff[s] = 0.0;
            if (c2 < 0 && typebc[m->nbc + c2] != kBuffer) ff[s] = 1.0;
        }
    });
}

int tflows_bfile_load(const char* path, int32_t hot, tflows_bfile* out) {
    return mesh_guard([&] {
        if (!out) throw std::invalid_argument("out is NULL");
        auto b = std::make_unique<tflows_bfile_s>();
        load_b(*b, path, hot != 0);
        *out = b.release();
    });
}

int tflows_bfile_destroy(tflows_bfile b) {
    return mesh_guard([&] { delete b; });
}

int tflows_bfile_info(tflows_bfile b, int32_t* nmat, int32_t* nb, int32_t* ninit, double* visc) {
    return mesh_guard([&] {
        if (!b) throw std::invalid_argument("null handle");
        if (nmat) *nmat = b->nmat;
        if (nb) *nb = b->nb;
        if (ninit) *ninit = b->ninit;
        if (visc) *visc = b->visc;
    });
}

int tflows_bfile_materials(tflows_bfile b, int32_t* statemat, double* denc, double* conc, double* capc) {
    return mesh_guard([&] {
        if (!b) throw std::invalid_argument("null handle");
        for (int n = 0; n < b->nmat; ++n) {
            if (statemat) statemat[n] = b->statemat[n];
            if (denc) denc[n] = b->denc[n];
            if (conc) conc[n] = b->conc[n];
            if (capc) capc[n] = b->capc[n];
        }
    });
}

int tflows_bfile_bounds(tflows_bfile b, int32_t* typbou, int32_t* from_file, double* u, double* v, double* w,
                        double* p, double* zmix) {
    return mesh_guard([&] {
        if (!b) throw std::invalid_argument("null handle");
        for (int n = 0; n < b->nb; ++n) {
            if (typbou) typbou[n] = b->typbou[n];
            if (from_file) from_file[n] = b->bfile[n];
            if (u) u[n] = b->bu[n];
            if (v) v[n] = b->bv[n];
            if (w) w[n] = b->bw[n];
            if (p) p[n] = b->bp[n];
            if (zmix) zmix[n] = b->bz[n];
        }
    });
}

int tflows_bfile_inits(tflows_bfile b, int32_t* from_file, double* u, double* v, double* w, double* zmix) {
    return mesh_guard([&] {
        if (!b) throw std::invalid_argument("null handle");
        for (int n = 0; n < b->ninit; ++n) {
            if (from_file) from_file[n] = b->ifile[n];
            if (u) u[n] = b->iu[n];
            if (v) v[n] = b->iv[n];
            if (w) w[n] = b->iw[n];
            if (zmix) zmix[n] = b->iz[n];
        }
    });
}

int tflows_bfile_name(tflows_bfile b, int32_t kind, int32_t n, char* buf, int32_t len) {
    return mesh_guard([&] {
        if (!b || !buf || len < 1) throw std::invalid_argument("null handle or buffer");
        const std::vector<std::string>& names = kind == 0 ? b->profile : b->initname;
        if (kind != 0 && kind != 1) throw std::invalid_argument("kind must be 0 (profile) or 1 (initial condition)");
        if (n < 1 || n > (int)names.size()) throw std::invalid_argument("entry out of range");
        const std::string& s = names[n - 1];
        if ((int)s.size() >= len) throw std::invalid_argument("buffer too short");
        std::memcpy(buf, s.c_str(), s.size() + 1);
    });
}

}  // extern "C"
