// C++ host side of the scan: parsing, set-up, interaction lowering, run,
// output.  See simulation.hpp.
#include "simulation.hpp"

#include <sys/stat.h>
#include <zlib.h>

#include <atomic>

#include <algorithm>
#include <array>
#include <initializer_list>
#include <cctype>
#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>

#include "../ctx_ref.hpp"
#include "cube_values.hpp"
#include "format.hpp"

namespace mafm {

namespace {

const double SIXTHRT2 = 1.12246204830937298142;   // globals.hpp:14
const double PI = 3.14159265358979323846;         // globals.hpp:13
const double FORCE_GRID_MARGIN = 1.5;             // globals.hpp:16
const double TIP_GAUSSIAN_WIDTH = 0.5;            // globals.hpp:18
const double HARTREE_TO_EV = 27.211386;           // globals.hpp:21-23
const double HARTREE_TO_KJ = 2625.5002;
const double HARTREE_TO_KCAL = 627.50961;
const double BOHR_TO_ANGST = 0.52917721092;       // cube_io.hpp:17

[[noreturn]] void fail(mafm_status s, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw Error(s, buf);
}

void check(mafm_ctx* ctx, mafm_status s, const char* what) {
    if (!ctx) fail(MAFM_ERR_STATE, "%s needs a device context (this simulation was opened without one)", what);
    if (s != MAFM_OK) fail(s, "%s: %s", what, mafm_last_error(ctx));
}

// checkForComments (parse.cpp:17-27); whitespace-only lines are skipped too.
bool skip_line(const std::string& line) {
    if (line.empty() || line[0] == '#' || line[0] == '%' || line[0] == '\n' || line[0] == '\r') return true;
    for (char ch : line)
        if (!std::isspace((unsigned char)ch)) return false;
    return true;
}

std::vector<std::string> split(const std::string& line) {
    std::istringstream is(line);
    std::vector<std::string> t;
    std::string w;
    while (is >> w) t.push_back(w);
    return t;
}

std::string lower(std::string s) {
    for (char& ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

double to_d(const std::string& s) { return std::strtod(s.c_str(), nullptr); }   // atof

// isint (utility.cpp:26-37)
bool is_int(const std::string& s) {
    for (char ch : s)
        if (!(std::isdigit((unsigned char)ch) || ch == '-' || ch == '+')) return false;
    return true;
}

bool on_off(const std::string& key, const std::string& v) {
    if (v == "on") return true;
    if (v == "off") return false;
    fail(MAFM_ERR_INVALID, "Option %s must be either on or off!", key.c_str());
}

// lines of a text file; a missing file is the reference's error `missing`
// (a printf format taking the path)
std::vector<std::string> read_lines(const std::string& path, const char* missing) {
    std::ifstream f(path);
    if (!f) fail(MAFM_ERR_INVALID, missing, path.c_str());
    std::vector<std::string> lines;
    std::string l;
    while (std::getline(f, l)) lines.push_back(l);
    return lines;
}

// Vector3 arithmetic of vectors.hpp:99-122: len() = sqrt(x^2 + y^2 + z^2)
// summed left to right, normalized() with its 1e-10 cut-off.
double vlen(const Vec3d& v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
double vdot(const Vec3d& a, const Vec3d& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
Vec3d vcross(const Vec3d& a, const Vec3d& b) {
    return Vec3d(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
Vec3d vnormalized(const Vec3d& v) {
    const double l = vlen(v);
    return l > 1e-10 ? Vec3d(v.x / l, v.y / l, v.z / l) : Vec3d();
}

bool same_pair(const std::string& a, const std::string& b, const std::string& c, const std::string& d) {
    return (a == c && b == d) || (a == d && b == c);
}

mafm_flex_term flex_term(int kind, std::initializer_list<int> atoms, std::initializer_list<double> p,
                         Vec3d shift = Vec3d()) {
    mafm_flex_term t;
    memset(&t, 0, sizeof t);
    t.kind = kind;
    int k = 0;
    for (int a : atoms) t.atom[k++] = a;
    k = 0;
    for (double v : p) t.p[k++] = v;
    t.shift[0] = shift.x;
    t.shift[1] = shift.y;
    t.shift[2] = shift.z;
    return t;
}

}  // namespace

// ---------------------------------------------------------------- Mat3d
double Mat3d::determinant() const {   // matrices.hpp:135-143
    double det = m[0] * (m[4] * m[8] - m[5] * m[7]);
    det += m[1] * (m[5] * m[6] - m[3] * m[8]);
    det += m[2] * (m[3] * m[7] - m[4] * m[6]);
    return det;
}

Mat3d Mat3d::inverse() const {        // matrices.hpp:89-114
    Mat3d o;
    double inv_det = 1 / determinant();
    double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7], i = m[8];
    o.m[0] = inv_det * (e * i - f * h);
    o.m[1] = -inv_det * (b * i - c * h);
    o.m[2] = inv_det * (b * f - c * e);
    o.m[3] = -inv_det * (d * i - f * g);
    o.m[4] = inv_det * (a * i - c * g);
    o.m[5] = -inv_det * (a * f - c * d);
    o.m[6] = inv_det * (d * h - e * g);
    o.m[7] = -inv_det * (a * h - b * g);
    o.m[8] = inv_det * (a * e - b * d);
    return o;
}

bool Mat3d::isDiagonal() const {      // matrices.hpp:72-86
    double diag = 0, nondiag = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) (i == j ? diag : nondiag) += std::fabs(m[3 * i + j]);
    return nondiag < 1.0e-10 * diag;
}

// ---------------------------------------------------------------- lowering
// Tip-surface pair terms become sites in list order; a Coulomb term folds
// into the site of the same atom that precedes it (the reference emits
// VDW(1,i) then Coulomb(1,i), simulation.cpp:309-312).
void LJInteraction::lower(Lowering& out) const {
    if (i1_ == 1) {
        const Vec3d& p = (*out.positions)[i2_];
        mafm_site s{};
        s.x = p.x + shift_.x; s.y = p.y + shift_.y; s.z = p.z + shift_.z;
        s.vdw = MAFM_VDW_LJ;
        s.p0 = es6_; s.p1 = es12_;
        out.sites.push_back(s);
        out.site_atom.push_back(i2_);
        out.site_shift.push_back(shift_);
    } else if (i1_ == 0 && i2_ == 1) {
        out.tip_dummy.vdw = MAFM_VDW_LJ;
        out.tip_dummy.p0 = es6_; out.tip_dummy.p1 = es12_; out.tip_dummy.p2 = 0;
        out.tip_dummy_vdw_set = true;
    } else {
        fail(MAFM_ERR_INVALID, "surface-surface interactions need flexible mode, which is not built");
    }
}

void MorseInteraction::lower(Lowering& out) const {
    if (i1_ == 1) {
        const Vec3d& p = (*out.positions)[i2_];
        mafm_site s{};
        s.x = p.x + shift_.x; s.y = p.y + shift_.y; s.z = p.z + shift_.z;
        s.vdw = MAFM_VDW_MORSE;
        s.p0 = de_; s.p1 = a_; s.p2 = re_;
        out.sites.push_back(s);
        out.site_atom.push_back(i2_);
        out.site_shift.push_back(shift_);
    } else if (i1_ == 0 && i2_ == 1) {
        out.tip_dummy.vdw = MAFM_VDW_MORSE;
        out.tip_dummy.p0 = de_; out.tip_dummy.p1 = a_; out.tip_dummy.p2 = re_;
        out.tip_dummy_vdw_set = true;
    } else {
        fail(MAFM_ERR_INVALID, "surface-surface interactions need flexible mode, which is not built");
    }
}

void CoulombInteraction::lower(Lowering& out) const {
    if (i1_ == 1) {
        if (out.sites.empty()) fail(MAFM_ERR_INVALID, "Coulomb term without a preceding VDW site");
        mafm_site& s = out.sites.back();
        const Vec3d& p = (*out.positions)[i2_];
        if (s.x != p.x || s.y != p.y || s.z != p.z || s.has_coulomb)
            fail(MAFM_ERR_INVALID, "Coulomb term does not follow the VDW term of its atom");
        s.has_coulomb = 1;
        s.qq = qq_;
    } else if (i1_ == 0 && i2_ == 1) {
        out.tip_dummy.has_coulomb = 1;
        out.tip_dummy.qq = qq_;
    } else {
        fail(MAFM_ERR_INVALID, "surface-surface interactions need flexible mode, which is not built");
    }
}

// flexible on: each Interaction becomes one term of the device list
void LJInteraction::lowerFlex(std::vector<mafm_flex_term>& out) const {
    out.push_back(flex_term(MAFM_FX_LJ, {i1_, i2_}, {es6_, es12_}, shift_));
}

void MorseInteraction::lowerFlex(std::vector<mafm_flex_term>& out) const {
    out.push_back(flex_term(MAFM_FX_MORSE, {i1_, i2_}, {de_, a_, re_}, shift_));
}

void CoulombInteraction::lowerFlex(std::vector<mafm_flex_term>& out) const {
    out.push_back(flex_term(MAFM_FX_COULOMB, {i1_, i2_}, {qq_}));
}

void TipHarmonicInteraction::lowerFlex(std::vector<mafm_flex_term>& out) const {
    out.push_back(flex_term(MAFM_FX_TIP_HARMONIC, {i1_, i2_}, {k_, r0_}));
}

void GridInteraction::lowerFlex(std::vector<mafm_flex_term>&) const {
    fail(MAFM_ERR_INVALID, "Cannot use a flexible molecule with a static force grid!");
}

void ElectrostaticPotentialInteraction::lowerFlex(std::vector<mafm_flex_term>&) const {
    fail(MAFM_ERR_INVALID, "External potential can be used only for non-flexible systems!");
}

void FlexOnlyInteraction::lower(Lowering&) const {
    fail(MAFM_ERR_INVALID, "bonded and substrate terms exist only with flexible on");
}

void XYHarmonicInteraction::lowerFlex(std::vector<mafm_flex_term>& out) const {
    out.push_back(flex_term(MAFM_FX_XY_HARMONIC, {i_}, {k_, p0_.x, p0_.y}));
}

SubstrateInteraction::SubstrateInteraction(int i, double eps, double sig, double rc, double lambda)
    : i_(i), sig_(sig), rc_(rc) {   // interactions.hpp:222-229
    multiplier_ = 2 * PI * eps * std::pow(sig / lambda, 2);
    ulj_ = (PI / (lambda * lambda)) * 4 * eps * (std::pow(sig / rc, 12) - std::pow(sig / rc, 6));
    ushift_ = multiplier_ * ((12. / 5) * std::pow(sig / rc, 10) - 3 * std::pow(sig / rc, 4));
    z0_ = -sig;
}

void SubstrateInteraction::lowerFlex(std::vector<mafm_flex_term>& out) const {
    out.push_back(flex_term(MAFM_FX_SUBSTRATE, {i_}, {multiplier_, ulj_, ushift_, z0_, rc_, sig_}));
}

void HarmonicInteraction::lowerFlex(std::vector<mafm_flex_term>& out) const {
    out.push_back(flex_term(MAFM_FX_BOND, {i1_, i2_}, {k_, r0_}));
}

void HarmonicAngleInteraction::lowerFlex(std::vector<mafm_flex_term>& out) const {
    out.push_back(flex_term(MAFM_FX_ANGLE, {shared_, i1_, i2_}, {k_, theta0_}));
}

void HarmonicDihedralInteraction::lowerFlex(std::vector<mafm_flex_term>& out) const {
    out.push_back(flex_term(MAFM_FX_DIHEDRAL, {i1_, i2_, i3_, i4_}, {k_, sigma0_}));
}

void factorImages(Lowering& low) {
    const size_t n = low.sites.size();
    std::vector<Vec3d> shifts;
    for (const Vec3d& v : low.site_shift) {
        bool seen = false;
        for (const Vec3d& u : shifts) seen |= (u == v);
        if (!seen) shifts.push_back(v);
    }
    if (shifts.size() <= 1 || n % shifts.size() != 0) return;
    const size_t block = n / shifts.size();
    for (size_t k = 0; k < n; ++k) {
        if (!(low.site_shift[k] == shifts[k / block]) || low.site_atom[k] != low.site_atom[k % block]) return;
        if (low.sites[k].has_coulomb) return;
    }
    std::vector<mafm_site> base(low.sites.begin(), low.sites.begin() + (long)block);
    for (size_t k = 0; k < block; ++k) {
        const Vec3d& p = (*low.positions)[low.site_atom[k]];
        base[k].x = p.x;
        base[k].y = p.y;
        base[k].z = p.z;
    }
    low.sites.swap(base);
    low.site_atom.resize(block);
    low.site_shift.assign(block, Vec3d());
    low.image_shifts.clear();
    for (const Vec3d& v : shifts) {
        low.image_shifts.push_back(v.x);
        low.image_shifts.push_back(v.y);
        low.image_shifts.push_back(v.z);
    }
    low.n_images = (int)shifts.size();
}

void TipHarmonicInteraction::lower(Lowering& out) const {
    if (!(i1_ == 0 && i2_ == 1)) fail(MAFM_ERR_INVALID, "tip harmonic must join dummy (0) and tip (1)");
    out.tip_dummy.k = k_;
    out.tip_dummy.r0 = r0_;
}

// ---------------------------------------------------------------- System
void System::initialize(int n_atoms) {           // system.cpp:10-25
    tip_pbc_ = false;
    n_atoms_ = n_atoms + 2;
    positions_.assign(n_atoms_, Vec3d());
    charges_.assign(n_atoms_, 0);
    fixed_.assign(n_atoms_, 0);
    types_.assign(n_atoms_, "");
    fixed_[0] = 1;
    fixed_[1] = 0;
}

void System::setMoleculeZ() {                    // system.cpp:103-116
    double min_z = 10e10;
    for (int i = 2; i < n_atoms_; ++i)
        if (positions_[i].z < min_z && types_[i] != "H") min_z = positions_[i].z;
    for (int i = 2; i < n_atoms_; ++i) positions_[i].z -= min_z;
    offset_.z = -min_z;
}

void System::centerMolecule(Vec2d pos) {         // system.cpp:119-136
    double avgx = 0.0, avgy = 0.0;
    for (int i = 2; i < n_atoms_; ++i) {
        avgx += positions_[i].x;
        avgy += positions_[i].y;
    }
    avgx /= n_atoms_ - 2;
    avgy /= n_atoms_ - 2;
    double dx = pos.x - avgx, dy = pos.y - avgy;
    for (int i = 2; i < n_atoms_; ++i) {
        positions_[i].x += dx;
        positions_[i].y += dy;
    }
    offset_.x = dx;
    offset_.y = dy;
}

void System::setUnitCell(const std::vector<Vec3d>& v) {   // system.cpp:170-177
    for (int i = 0; i < 3; ++i) {
        cell_.at(0, i) = v[i].x;
        cell_.at(1, i) = v[i].y;
        cell_.at(2, i) = v[i].z;
    }
    tip_pbc_ = true;
}

void System::rotateCoordAxes(const std::string& seq) {    // system.cpp:68-90
    for (auto& p : positions_) {
        Vec3d q = p;
        if (seq == "ZXY") p = Vec3d(q.z, q.x, q.y);
        else if (seq == "YZX") p = Vec3d(q.y, q.z, q.x);
        else fail(MAFM_ERR_INVALID, "Cannot rotate coordinate axes to %s!", seq.c_str());
    }
}

// ---------------------------------------------------------------- cube
Cube readCube(const std::string& path, bool with_values) {   // cube_io.cpp:10-139
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) fail(MAFM_ERR_INVALID, "Could not open cube file %s", path.c_str());
    std::string buf;
    {
        fseek(f, 0, SEEK_END);
        long len = ftell(f);
        fseek(f, 0, SEEK_SET);
        buf.resize((size_t)std::max(len, 0L));
        size_t got = len > 0 ? fread(&buf[0], 1, (size_t)len, f) : 0;
        fclose(f);
        buf.resize(got);
    }
    const char* p = buf.c_str();
    const char* end = p + buf.size();
    for (int l = 0; l < 2; ++l) {                 // two comment lines
        while (p < end && *p != '\n') ++p;
        if (p < end) ++p;
    }
    auto num = [&](const char* what) -> double {
        char* q = nullptr;
        errno = 0;
        double v = std::strtod(p, &q);
        if (q == p) fail(MAFM_ERR_INVALID, "Could not read %s from the cube file. Check the format of the file.", what);
        p = q;
        return v;
    };
    Cube c;
    c.n_atoms = (int)num("atom count");
    c.origin.x = num("origin"); c.origin.y = num("origin"); c.origin.z = num("origin");
    for (int a = 0; a < 3; ++a) {
        c.n[a] = (int)num("voxel count");
        c.voxel[a].x = num("voxel vector");
        c.voxel[a].y = num("voxel vector");
        c.voxel[a].z = num("voxel vector");
    }
    if (c.n[0] < 0) {                            // Angstrom
        for (int a = 0; a < 3; ++a) c.n[a] = -c.n[a];
    } else {                                     // Bohr
        for (int a = 0; a < 3; ++a) c.voxel[a] = BOHR_TO_ANGST * c.voxel[a];
        c.origin = BOHR_TO_ANGST * c.origin;
    }
    for (int a = 0; a < std::abs(c.n_atoms); ++a)
        for (int q = 0; q < 5; ++q) num("atom record");
    if (with_values) {
        const size_t total = (size_t)c.n[0] * c.n[1] * c.n[2];
        if (!parse_cube_values(p, end, total, c.values))
            fail(MAFM_ERR_INVALID, "Could not read the volumetric data from the cube file. Check the format of the file.");
    }
    return c;
}

// ---------------------------------------------------------------- Simulation
Simulation::~Simulation() {
    for (int h : owned_grids_) (void)mafm_grid_free(ctx_, h);
}

void Simulation::parseCommandLine(const std::string& input, const std::string& output) {   // parse.cpp:30-80
    InputOptions& o = options_;
    o.inputfile = input;
    size_t sl = input.rfind('/');
    o.inputfolder = sl != std::string::npos ? input.substr(0, sl + 1) : "";
    o.outputfolder = output;
    if (!o.outputfolder.empty() && o.outputfolder.back() != '/') o.outputfolder += '/';
    if (!o.outputfolder.empty()) {
        std::string acc;
        for (size_t i = 0; i < o.outputfolder.size(); ++i) {
            acc += o.outputfolder[i];
            if (o.outputfolder[i] == '/') mkdir(acc.c_str(), 0755);
        }
    }
}

void Simulation::readInputFile() {                // parse.cpp:83-405
    InputOptions& o = options_;
    for (const std::string& line : read_lines(o.inputfile, "The file %s does not exist!")) {
        if (skip_line(line)) continue;
        std::vector<std::string> t = split(line);
        const std::string key = lower(t[0]);
        const std::string val = t.size() > 1 ? t[1] : "";
        auto need = [&](size_t n) {
            if (t.size() < n) fail(MAFM_ERR_INVALID, "Option %s needs %zu values", key.c_str(), n - 1);
        };
        if (key == "xyzfile") o.xyzfile = o.inputfolder + val;
        else if (key == "paramfile") o.paramfile = o.inputfolder + val;
        else if (key == "e_potential_file") o.e_potential_file = o.inputfolder + val;
        else if (key == "tipatom") o.tipatom = val;
        else if (key == "dummyatom") o.dummyatom = val;
        else if (key == "area") { need(3); o.area = {to_d(t[1]), to_d(t[2])}; }
        else if (key == "center") { need(3); o.center = {to_d(t[1]), to_d(t[2])}; }
        else if (key == "zhigh") o.zhigh = to_d(val);
        else if (key == "zlow") o.zlow = to_d(val);
        else if (key == "dx") o.dx = to_d(val);
        else if (key == "dy") o.dy = to_d(val);
        else if (key == "dz") o.dz = to_d(val);
        else if (key == "cell_a") { need(4); o.cell_a = Vec3d(to_d(t[1]), to_d(t[2]), to_d(t[3])); }
        else if (key == "cell_b") { need(4); o.cell_b = Vec3d(to_d(t[1]), to_d(t[2]), to_d(t[3])); }
        else if (key == "cell_c") { need(4); o.cell_c = Vec3d(to_d(t[1]), to_d(t[2]), to_d(t[3])); }
        else if (key == "etol") o.etol = to_d(val);
        else if (key == "ftol") o.ftol = to_d(val);
        else if (key == "dt") o.dt = to_d(val);
        else if (key == "maxsteps") o.maxsteps = std::atoi(val.c_str());
        else if (key == "bufsize") o.bufsize = std::atoi(val.c_str());
        else if (key == "gzip") o.gzip = on_off(key, val);
        else if (key == "statistics") o.statistics = on_off(key, val);
        else if (key == "flexible") o.flexible = on_off(key, val);
        else if (key == "rigidgrid") o.rigidgrid = on_off(key, val);
        else if (key == "coulomb") o.coulomb = on_off(key, val);
        else if (key == "tip_dummy_coulomb") o.tip_dummy_coulomb = on_off(key, val);
        else if (key == "use_external_potential") o.use_external_potential = on_off(key, val);
        else if (key == "vdw_pbc") o.vdw_pbc = on_off(key, val);
        else if (key == "surface_normal") {
            std::string v = lower(val);
            o.normal_text = val;
            if (v == "z") o.normal = NORMAL_Z;
            else if (v == "y") o.normal = NORMAL_Y;
            else if (v == "x") o.normal = NORMAL_X;
            else fail(MAFM_ERR_INVALID, "Option %s must be either X, Y or Z!", key.c_str());
        } else if (key == "minterm") {
            o.minterm_text = val;
            if (val == "e") o.minterm = MIN_E;
            else if (val == "f") o.minterm = MIN_F;
            else if (val == "ef") o.minterm = MIN_EF;
            else fail(MAFM_ERR_INVALID, "Option %s must be either e, f or ef!", key.c_str());
        } else if (key == "units") {
            InteractionParameters& ip = interaction_parameters_;
            o.units_text = val;
            if (val == "kcal/mol" || val == "kcal") { o.units = U_KCAL; ip.qbase = 332.06371; }
            else if (val == "kJ/mol" || val == "kJ") { o.units = U_KJ; ip.qbase = 1389.354563; }
            else if (val == "eV") { o.units = U_EV; ip.qbase = 14.39964901; }
            else fail(MAFM_ERR_INVALID, "Option %s must be either kcal/mol (default), kJ/mol or eV!", key.c_str());
            o.units_given = true;
            ip.qbase_set = true;
        } else if (key == "minimiser") {
            if (val == "SD") o.minimiser_type = STEEPEST_DESCENT;
            else if (val == "FIRE") o.minimiser_type = FIRE;
            else fail(MAFM_ERR_INVALID, "Unrecognised minimiser type!");
        } else if (key == "integrator") {
            if (val == "euler") o.integrator_type = EULER;
            else if (val == "midpoint") o.integrator_type = MIDPOINT;
            else if (val == "rk4") o.integrator_type = RK4;
            else fail(MAFM_ERR_INVALID, "Unrecognised integrator type!");
        } else {
            fail(MAFM_ERR_INVALID, "Unknown option %s!", key.c_str());
        }
    }
    if (o.xyzfile.empty()) fail(MAFM_ERR_INVALID, "Specify at least an xyzfile!");
    if (o.paramfile.empty()) fail(MAFM_ERR_INVALID, "Specify at least a parameter file!");
    if (o.tipatom.empty()) fail(MAFM_ERR_INVALID, "Specify at least a tip atom!");
    if (o.dummyatom.empty()) fail(MAFM_ERR_INVALID, "Specify at least a dummy atom!");
    if (o.minterm == NOT_SET)
        fail(MAFM_ERR_INVALID, "Specify at least a minimization termination criterion (e, f, or ef)!");
    if (o.center.x == -1 && o.center.y == -1) o.center = {o.area.x / 2, o.area.y / 2};
    if (o.rigidgrid && o.flexible) fail(MAFM_ERR_INVALID, "Cannot use a flexible molecule with a static force grid!");
    if (o.coulomb && o.use_external_potential)
        fail(MAFM_ERR_INVALID, "Cannot use Coulomb interaction and external electrostatic potential at the same time!");
    if (o.use_external_potential && o.flexible)
        fail(MAFM_ERR_INVALID, "External potential can be used only for non-flexible systems!");
    if (o.use_external_potential && o.e_potential_file.empty())
        fail(MAFM_ERR_INVALID, "If you want to use external electrostatic potential, you must specify a file that contains it!");
    if (o.vdw_pbc && o.coulomb)
        fail(MAFM_ERR_INVALID, "Implementation of Coulomb interaction does not support any periodic boundary conditions!");
    if (o.vdw_pbc && !o.use_external_potential &&
        (o.cell_a == Vec3d() || o.cell_b == Vec3d() || o.cell_c == Vec3d()))
        fail(MAFM_ERR_INVALID, "The unit cell vectors must be given if periodic vdW is used.");
    if (!(o.dx > 0 && o.dy > 0 && o.dz > 0)) fail(MAFM_ERR_INVALID, "dx, dy and dz must be positive");
    // openUniverse (mechafm.cpp:47-49)
    n_points_[0] = (int)std::floor(o.area.x / o.dx) + 1;
    n_points_[1] = (int)std::floor(o.area.y / o.dy) + 1;
    n_points_[2] = (int)std::floor((o.zhigh - o.zlow) / o.dz) + 1;
    if (n_points_[0] <= 0 || n_points_[1] <= 0 || n_points_[2] <= 0)
        fail(MAFM_ERR_INVALID, "empty scan grid %d x %d x %d", n_points_[0], n_points_[1], n_points_[2]);
}

void Simulation::readXYZFile() {                  // parse.cpp:477-585
    InputOptions& o = options_;
    std::vector<std::string> lines = read_lines(o.xyzfile, "No such file: %s!");
    int n_atoms = 0;
    bool first = true, realxyz = false;
    for (const std::string& line : lines) {
        if (skip_line(line)) continue;
        if (first) {
            std::vector<std::string> t = split(line);
            if (is_int(t[0])) {
                n_atoms = std::atoi(t[0].c_str());
                realxyz = true;
                break;
            }
            first = false;
        }
        ++n_atoms;
    }
    if (n_atoms <= 0) fail(MAFM_ERR_INVALID, "no atoms in %s", o.xyzfile.c_str());
    system.initialize(n_atoms);
    system.types_[0] = o.dummyatom;
    system.types_[1] = o.tipatom;
    int atom_i = 2, ncols = 0;
    for (size_t n = 0; n < lines.size(); ++n) {
        if (realxyz && n < 2) continue;
        if (skip_line(lines[n])) continue;
        std::vector<std::string> t = split(lines[n]);
        if (ncols == 0) ncols = (int)t.size();
        if (ncols < 4 || ncols > 6 || (int)t.size() < ncols)
            fail(MAFM_ERR_INVALID, "Invalid number of columns in the xyz file.");
        if (atom_i >= system.n_atoms_) fail(MAFM_ERR_INVALID, "more atoms in %s than declared", o.xyzfile.c_str());
        system.types_[atom_i] = t[0];
        system.positions_[atom_i] = Vec3d(to_d(t[1]), to_d(t[2]), to_d(t[3]));
        system.charges_[atom_i] = ncols >= 5 ? to_d(t[4]) : 0.0;
        // the sixth column when flexible, else every surface atom fixed (parse.cpp:548-563)
        system.fixed_[atom_i] = o.flexible ? (ncols == 6 ? std::atoi(t[5].c_str()) : 0) : 1;
        ++atom_i;
    }
    if (atom_i != system.n_atoms_) fail(MAFM_ERR_INVALID, "fewer atoms in %s than declared", o.xyzfile.c_str());
    double charge_check = 0;
    for (int i = 0; i < system.n_atoms_; ++i) charge_check += std::pow(system.charges_[i], 2);
    o.xyz_charges = !(std::fabs(charge_check) < 1e-10);
    if (!o.xyz_charges)                            // parse.cpp:575-581
        warnings_.push_back("The RMSQ-error for the charges read from the XYZ-file is zero.\n"
                            "+-     Charges will be read from the parameter file.\n"
                            "+-     If you want zero charge, set the charge in the parameter file to zero.\n"
                            "+-");
}

std::string Simulation::settingsText() const {
    const InputOptions& o = options_;
    std::string s;
    for (const std::string& w : warnings_) s += "+- WARNING: " + w + "\n";
    auto put = [&s](const char* fmt, ...) {
        char line[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(line, sizeof line, fmt, ap);
        va_end(ap);
        s += "+- ";
        s += line;
        s += "\n";
    };
    auto onoff = [](bool b) { return b ? "on" : "off"; };
    put("");
    put("Input settings for        %s:", o.inputfile.c_str());
    put("");
    put("xyzfile:                  %-s", o.xyzfile.c_str());
    put("paramfile:                %-s", o.paramfile.c_str());
    put("tipatom:                  %-s", o.tipatom.c_str());
    put("dummyatom:                %-s", o.dummyatom.c_str());
    put("");
    put("units:                    %-s", o.units_text.c_str());
    put("");
    put("minterm:                  %-s", o.minterm_text.c_str());
    put("etol:                     %-8.4f", o.etol);
    put("ftol:                     %-8.4f", o.ftol);
    put("dt:                       %-8.4f", o.dt);
    put("maxsteps:                 %-8d", o.maxsteps);
    put("");
    put("area:                     %-8.4f %-8.4f", o.area.x, o.area.y);
    put("center:                   %-8.4f %-8.4f", o.center.x, o.center.y);
    put("zhigh:                    %-8.4f", o.zhigh);
    put("zlow:                     %-8.4f", o.zlow);
    put("dx:                       %-8.4f", o.dx);
    put("dy:                       %-8.4f", o.dy);
    put("dz:                       %-8.4f", o.dz);
    put("");
    put("surface_normal:           %-s", o.normal_text.c_str());
    put("vdw_pbc:                  %-s", onoff(o.vdw_pbc));
    if (o.vdw_pbc && !o.use_external_potential) {
        put("cell_a:                   %-8.4f %-8.4f %-8.4f", o.cell_a.x, o.cell_a.y, o.cell_a.z);
        put("cell_b:                   %-8.4f %-8.4f %-8.4f", o.cell_b.x, o.cell_b.y, o.cell_b.z);
        put("cell_c:                   %-8.4f %-8.4f %-8.4f", o.cell_c.x, o.cell_c.y, o.cell_c.z);
    }
    put("");
    put("coulomb:                  %-s", onoff(o.coulomb));
    put("tip_dummy_coulomb:        %-s", onoff(o.tip_dummy_coulomb));
    put("use_external_potential:   %-s", onoff(o.use_external_potential));
    if (o.use_external_potential) put("e_potential_file:         %-s", o.e_potential_file.c_str());
    put("");
    put("flexible:                 %-s", onoff(o.flexible));
    put("rigidgrid:                %-s", onoff(o.rigidgrid));
    put("");
    put("minimiser:         %-s", o.minimiser_type == FIRE ? "FIRE" : "SD");
    put("integrator:        %-s", o.integrator_type == EULER ? "euler" : (o.integrator_type == RK4 ? "rk4" : "midpoint"));
    put("");
    put("bufsize:           %-8d", o.bufsize);
    put("gzip:              %-s", onoff(o.gzip));
    put("statistics:        %-s", onoff(o.statistics));
    put("");
    return s;
}

void Simulation::readParameterFile() {            // parse.cpp:588-697
    InputOptions& o = options_;
    InteractionParameters& p = interaction_parameters_;
    for (const std::string& line : read_lines(o.paramfile, "No parameter file %s found!")) {
        if (skip_line(line)) continue;
        std::vector<std::string> t = split(line);
        if (t[0] == "atom") {
            if (t.size() < 6) fail(MAFM_ERR_INVALID, "atom line needs name eps sig mass charge");
            AtomParameters& a = p.atom_parameters[t[1]];
            a.eps = to_d(t[2]);
            a.sig = to_d(t[3]);
            a.mass = to_d(t[4]);
            a.q = to_d(t[5]);
            if (o.dummyatom == t[1]) system.charges_[0] = a.q;
            if (o.tipatom == t[1]) system.charges_[1] = a.q;
        } else if (t[0] == "harm") {
            if (p.harm_set) fail(MAFM_ERR_INVALID, "Parameters for harmonic spring can only be specified once!");
            if (t.size() < 3) fail(MAFM_ERR_INVALID, "harm needs k and r0");
            p.tip_dummy_k = to_d(t[1]);
            p.tip_dummy_r0 = to_d(t[2]);
            p.harm_set = true;
        } else if (t[0] == "pair_ovwrt") {
            if (t.size() < 4) fail(MAFM_ERR_INVALID, "pair_ovwrt needs atom1 atom2 style parameters");
            OverwriteParameters ow;
            ow.a = t[1];
            ow.b = t[2];
            if (t[3] == "lj") {
                if (t.size() != 6) fail(MAFM_ERR_INVALID, "Only two parameters (eps, sig) allowed for LJ");
                ow.morse = false;
                ow.eps = to_d(t[4]);
                ow.sig = to_d(t[5]);
            } else if (t[3] == "morse") {
                if (t.size() != 7) fail(MAFM_ERR_INVALID, "Only three parameters (De, a, re) allowed for Morse");
                ow.morse = true;
                ow.de = to_d(t[4]);
                ow.alpha = to_d(t[5]);
                ow.re = to_d(t[6]);
            } else {
                fail(MAFM_ERR_INVALID, "Unknown pair style overwrite '%s'", t[3].c_str());
            }
            p.overwrite_parameters.push_back(ow);
        }
    }
    if (!p.harm_set) fail(MAFM_ERR_INVALID, "No harmonic spring parameters found in parameter file!");
    if (o.flexible) readFlexibleParameters(o.paramfile);
}

void Simulation::readFlexibleParameters(const std::string& path) {   // parse.cpp:700-782
    InteractionParameters& p = interaction_parameters_;
    bool bcheck = false, acheck = false, dcheck = false, scheck = false;
    for (const std::string& line : read_lines(path, "No parameter file %s found!")) {
        if (skip_line(line)) continue;
        std::vector<std::string> t = split(line);
        auto num = [&](size_t k, double& dst) {
            if (t.size() > k) dst = to_d(t[k]);
        };
        if (t[0] == "bond") {
            if (bcheck) warnings_.push_back("Parameters for harmonic bond defined multiple times!");
            num(1, p.bond_k);
            bcheck = true;
        } else if (t[0] == "angle") {
            if (acheck) warnings_.push_back("Parameters for harmonic angle defined multiple times!");
            num(1, p.angle_k);
            acheck = true;
        } else if (t[0] == "dihedral") {
            if (dcheck) warnings_.push_back("Parameters for harmonic dihedral defined multiple times!");
            num(1, p.dihedral_k);
            dcheck = true;
        } else if (t[0] == "substrate") {
            if (scheck) warnings_.push_back("Parameters for substrate support defined multiple times!");
            num(1, p.substrate_eps);
            num(2, p.substrate_sig);
            num(3, p.substrate_lambda);
            num(4, p.substrate_rc);
            num(5, p.substrate_k);
            scheck = true;
        } else if (t[0] == "topobond") {
            if (t.size() < 4) fail(MAFM_ERR_INVALID, "topobond needs atom1 atom2 and the bond length");
            PossibleBond pb;
            pb.a = t[1];
            pb.b = t[2];
            pb.r0 = to_d(t[3]);
            for (const PossibleBond& q : p.possible_bonds)
                if (same_pair(q.a, q.b, pb.a, pb.b)) {
                    warnings_.push_back("The topobond for " + pb.a + " and " + pb.b + " is defined at least twice!");
                    break;
                }
            p.possible_bonds.push_back(pb);
        }
    }
    if (!bcheck) fail(MAFM_ERR_INVALID, "No harmonic bond parameters found in parameter file!");
    if (!acheck) fail(MAFM_ERR_INVALID, "No harmonic angle parameters found in parameter file!");
    if (!dcheck) fail(MAFM_ERR_INVALID, "No harmonic dihedral parameters found in parameter file!");
    if (!scheck) fail(MAFM_ERR_INVALID, "No harmonic substrate support parameters found in parameter file!");
}

void Simulation::calculateTipDummyDistance() {    // simulation.cpp:218-225
    auto& ap = interaction_parameters_.atom_parameters;
    auto d = ap.find(system.types_[0]);
    auto t = ap.find(system.types_[1]);
    if (d == ap.end() || t == ap.end())
        fail(MAFM_ERR_INVALID, "Parameters for the tip or dummy atom type not found in parameter file!");
    system.setTipDummyDistance(mixsig(d->second.sig, t->second.sig) * SIXTHRT2);
}

bool Simulation::findOverwriteParameters(int i1, int i2, OverwriteParameters& op) const {   // :227-236
    const std::string& a = system.types_[i1];
    const std::string& b = system.types_[i2];
    for (const auto& ow : interaction_parameters_.overwrite_parameters) {
        if ((ow.a == a && ow.b == b) || (ow.a == b && ow.b == a)) {
            op = ow;
            return true;
        }
    }
    return false;
}

void Simulation::addVDWInteraction(int i1, int i2, Vec3d shift) {   // simulation.cpp:238-267
    OverwriteParameters op;
    if (findOverwriteParameters(i1, i2, op)) {
        if (op.morse) {
            interactions_.emplace_back(new MorseInteraction(i1, i2, op.de, op.alpha, op.re, shift));
        } else {
            double es6 = 4 * op.eps * std::pow(op.sig, 6);
            double es12 = 4 * op.eps * std::pow(op.sig, 12);
            interactions_.emplace_back(new LJInteraction(i1, i2, es6, es12, shift));
        }
        return;
    }
    auto& ap = interaction_parameters_.atom_parameters;
    auto a1 = ap.find(system.types_[i1]);
    auto a2 = ap.find(system.types_[i2]);
    if (a1 == ap.end())
        fail(MAFM_ERR_INVALID, "Parameters for atom type %s not found in parameter file!", system.types_[i1].c_str());
    if (a2 == ap.end())
        fail(MAFM_ERR_INVALID, "Parameters for atom type %s not found in parameter file!", system.types_[i2].c_str());
    double m_eps = mixeps(a1->second.eps, a2->second.eps);
    double m_sig = mixsig(a1->second.sig, a2->second.sig);
    double es6 = 4 * m_eps * std::pow(m_sig, 6);
    double es12 = 4 * m_eps * std::pow(m_sig, 12);
    interactions_.emplace_back(new LJInteraction(i1, i2, es6, es12, shift));
}

void Simulation::addCoulombInteraction(int i1, int i2) {   // simulation.cpp:269-295
    double q1, q2;
    if (options_.xyz_charges) {
        q1 = system.charges_[i1];
        q2 = system.charges_[i2];
    } else {
        auto& ap = interaction_parameters_.atom_parameters;
        auto a1 = ap.find(system.types_[i1]);
        auto a2 = ap.find(system.types_[i2]);
        if (a1 == ap.end() || a2 == ap.end())
            fail(MAFM_ERR_INVALID, "Parameters for atom type not found in parameter file!");
        q1 = a1->second.q;
        q2 = a2->second.q;
    }
    if (q1 != 0 && q2 != 0) {
        if (!interaction_parameters_.qbase_set)
            fail(MAFM_ERR_INVALID, "Coulomb terms need a `units` line (it sets the Coulomb prefactor)");
        double qq = interaction_parameters_.qbase * q1 * q2;
        interactions_.emplace_back(new CoulombInteraction(i1, i2, qq));
    }
}

// DataGrid::rotateCoordAxes("ZXY") (data_grid.cpp:49-70): new axes (z, x, y)
// of the old ones: counts, voxel vectors (rows and columns permuted, i.e.
// S^T B S), origin and the x-major / z-fastest values.
void rotateCubeZXY(Cube& c) {
    const int o[3] = {c.n[0], c.n[1], c.n[2]};
    const int perm[3] = {2, 0, 1};   // new axis a = old axis perm[a]
    Vec3d vox[3];
    for (int a = 0; a < 3; ++a) {
        const Vec3d& v = c.voxel[perm[a]];
        vox[a] = Vec3d(v.z, v.x, v.y);
    }
    const int n[3] = {o[2], o[0], o[1]};
    std::vector<double> out(c.values.size());
    for (int ix = 0; ix < n[0]; ++ix)
        for (int iy = 0; iy < n[1]; ++iy)
            for (int iz = 0; iz < n[2]; ++iz)
                out[((size_t)ix * n[1] + iy) * n[2] + iz] = c.values[((size_t)iy * o[1] + iz) * o[2] + ix];
    c.values.swap(out);
    for (int a = 0; a < 3; ++a) {
        c.n[a] = n[a];
        c.voxel[a] = vox[a];
    }
    c.origin = Vec3d(c.origin.z, c.origin.x, c.origin.y);
}

ForceGrid Simulation::buildElectrostaticPotential() {      // simulation.cpp:323-417
    Cube cube = readCube(options_.e_potential_file, true);
    double scale;
    if (options_.units == U_EV) scale = -HARTREE_TO_EV;
    else if (options_.units == U_KJ) scale = -HARTREE_TO_KJ;
    else scale = -HARTREE_TO_KCAL;
    for (double& v : cube.values) v *= scale;
    // surface_normal y: the grid takes the system's ZXY axis order
    // (simulation.cpp:358-362, data_grid.cpp:49-70); for x the reference asks
    // for ZYX, which DataGrid::rotateCoordAxes does not implement (it throws).
    if (options_.normal == NORMAL_X)
        fail(MAFM_ERR_INVALID, "an external potential with surface_normal x needs the grid axis order ZYX, which "
                               "the reference does not implement");
    if (options_.normal == NORMAL_Y) rotateCubeZXY(cube);
    mafm_potential_desc pd{};
    ForceGrid g;
    for (int a = 0; a < 3; ++a) {
        pd.n[a] = cube.n[a];
        g.n[a] = cube.n[a];
        g.basis.at(0, a) = cube.voxel[a].x;
        g.basis.at(1, a) = cube.voxel[a].y;
        g.basis.at(2, a) = cube.voxel[a].z;
    }
    memcpy(pd.basis, g.basis.m, sizeof pd.basis);
    g.offset = cube.origin + system.getOffset();
    pd.origin[0] = g.offset.x; pd.origin[1] = g.offset.y; pd.origin[2] = g.offset.z;
    pd.potential = cube.values.data();
    pd.tip_charge = system.charges_[1];
    pd.gaussian_width = TIP_GAUSSIAN_WIDTH;
    g.periodic = true;
    check(ctx_, mafm_potential_grid(ctx_, &pd, &g.handle), "external potential");
    owned_grids_.push_back(g.handle);
    return g;
}

void Simulation::buildTipSurfaceInteractions() {  // simulation.cpp:297-418
    if (options_.vdw_pbc) {
        const Mat3d& cell = system.getUnitCell();
        for (int a = -1; a <= 1; ++a)
            for (int b = -1; b <= 1; ++b) {
                Vec3d shift = (double)a * cell.column(0) + (double)b * cell.column(1);
                for (int i = 2; i < system.n_atoms_; ++i) addVDWInteraction(1, i, shift);
            }
    } else {
        for (int i = 2; i < system.n_atoms_; ++i) {
            addVDWInteraction(1, i);
            if (options_.coulomb) addCoulombInteraction(1, i);
        }
    }
    if (options_.use_external_potential) {
        ForceGrid g = buildElectrostaticPotential();
        interactions_.emplace_back(new ElectrostaticPotentialInteraction(g));
    }
}

void Simulation::buildTipGridInteractions() {     // simulation.cpp:420-501
    if (!interactions_.empty()) fail(MAFM_ERR_STATE, "Interaction list is not empty when building force grid!");
    buildTipSurfaceInteractions();
    const InputOptions& o = options_;
    const double sp[3] = {o.dx / 2, o.dy / 2, o.dz / 2};
    int border[3], n[3];
    for (int a = 0; a < 3; ++a) border[a] = (int)std::ceil(FORCE_GRID_MARGIN / sp[a]);
    n[0] = (int)std::floor(o.area.x / sp[0]) + 2 * border[0] + 1;
    n[1] = (int)std::floor(o.area.y / sp[1]) + 2 * border[1] + 1;
    n[2] = (int)std::floor((o.zhigh - o.zlow) / sp[2]) + 2 * border[2] + 1;
    Vec3d offset(-border[0] * sp[0], -border[1] * sp[1], o.zlow - system.getTipDummyDistance() - border[2] * sp[2]);
    // lower the tip-surface list (sites + optional potential grid) and sample it on the device
    Lowering low;
    low.positions = &system.positions_;
    for (const auto& it : interactions_) it->lower(low);
    factorImages(low);
    mafm_system_desc sd{};
    sd.n_sites = (int32_t)low.sites.size();
    sd.sites = low.sites.data();
    sd.n_images = low.n_images;
    sd.image_shifts = low.image_shifts.empty() ? nullptr : low.image_shifts.data();
    sd.force_grid = -1;
    sd.potential_grid = low.potential_grid;
    sd.tip_pbc = system.tipPbc() ? 1 : 0;
    memcpy(sd.cell, system.getUnitCell().m, sizeof sd.cell);
    sd.offset[0] = system.getOffset().x; sd.offset[1] = system.getOffset().y; sd.offset[2] = system.getOffset().z;
    mafm_grid_desc gd{};
    ForceGrid fg;
    for (int a = 0; a < 3; ++a) {
        gd.n[a] = n[a];
        fg.n[a] = n[a];
        force_grid_n_[a] = n[a];
        fg.basis.at(a, a) = sp[a];
    }
    memcpy(gd.basis, fg.basis.m, sizeof gd.basis);
    gd.offset[0] = offset.x; gd.offset[1] = offset.y; gd.offset[2] = offset.z;
    gd.periodic = 0;
    fg.offset = offset;
    fg.periodic = false;
    check(ctx_, mafm_grid_build(ctx_, &sd, &gd, &fg.handle), "force grid build");
    owned_grids_.push_back(fg.handle);
    interactions_.clear();
    interactions_.emplace_back(new GridInteraction(fg));
}

void Simulation::buildTipDummyInteractions() {    // simulation.cpp:503-515
    addVDWInteraction(0, 1);
    if (options_.tip_dummy_coulomb) addCoulombInteraction(0, 1);
    interactions_.emplace_back(new TipHarmonicInteraction(0, 1, interaction_parameters_.tip_dummy_k,
                                                          interaction_parameters_.tip_dummy_r0));
}

void Simulation::buildInteractions() {            // simulation.cpp:201-216
    interactions_.clear();
    if (options_.rigidgrid) buildTipGridInteractions();
    else buildTipSurfaceInteractions();
    buildTipDummyInteractions();
    if (options_.flexible) {
        buildSurfaceSurfaceInteractions();
        buildSubstrateInteractions();
    }
}

void Simulation::buildSubstrateInteractions() {   // simulation.cpp:517-531
    const InteractionParameters& ip = interaction_parameters_;
    for (int i = 2; i < system.n_atoms_; ++i) {
        interactions_.emplace_back(
            new SubstrateInteraction(i, ip.substrate_eps, ip.substrate_sig, ip.substrate_rc, ip.substrate_lambda));
        if (system.fixed_[i] == 2) {
            Vec2d p0;
            p0.x = system.positions_[i].x;
            p0.y = system.positions_[i].y;
            interactions_.emplace_back(new XYHarmonicInteraction(i, ip.substrate_k, p0));
        }
    }
}

// Bonds between atoms closer than 1.1 x a listed bond length, the angles of
// bond pairs sharing an atom, the dihedrals of angle pairs, and non-bonded
// VDW (+ Coulomb) pairs between atoms that no chain of bonds connects, in the
// reference's enumeration order (simulation.cpp:557-693); every equilibrium
// value is the input geometry's own.
void Simulation::buildSurfaceSurfaceInteractions() {
    const int n = system.n_atoms_;
    const std::vector<Vec3d>& P = system.positions_;
    const InteractionParameters& ip = interaction_parameters_;
    std::vector<std::vector<int>> adjacent(n);
    std::vector<std::pair<int, int>> bonds;
    for (int i = 2; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double r0 = 0;   // the last matching topobond
            for (const PossibleBond& pb : ip.possible_bonds)
                if (same_pair(pb.a, pb.b, system.types_[i], system.types_[j])) r0 = pb.r0;
            const double d = vlen(P[i] - P[j]);
            if (d < 1.1 * r0) {
                bonds.emplace_back(i, j);
                adjacent[i].push_back(j);
                adjacent[j].push_back(i);
                interactions_.emplace_back(new HarmonicInteraction(i, j, ip.bond_k, d));
            }
        }
    // getConnectedAtoms (simulation.cpp:533-555) with distance = n_atoms: every
    // atom reachable through bonds; an atom without bonds is connected to none
    std::vector<int> component(n, -1);
    for (int i = 0; i < n; ++i) {
        if (component[i] >= 0 || adjacent[i].empty()) continue;
        std::vector<int> stack{i};
        component[i] = i;
        while (!stack.empty()) {
            const int a = stack.back();
            stack.pop_back();
            for (int b : adjacent[a])
                if (component[b] < 0) {
                    component[b] = i;
                    stack.push_back(b);
                }
        }
    }
    std::vector<std::array<int, 3>> angles;   // shared, atom1, atom2
    for (size_t i = 0; i < bonds.size(); ++i)
        for (size_t j = i + 1; j < bonds.size(); ++j) {
            const std::pair<int, int> b1 = bonds[i], b2 = bonds[j];
            int shared = -1, a1 = 0, a2 = 0;
            if (b1.first == b2.first) {
                shared = b1.first; a1 = b1.second; a2 = b2.second;
            } else if (b1.first == b2.second) {
                shared = b1.first; a1 = b1.second; a2 = b2.first;
            } else if (b1.second == b2.first) {
                shared = b1.second; a1 = b1.first; a2 = b2.second;
            } else if (b1.second == b2.second) {
                shared = b1.second; a1 = b1.first; a2 = b2.first;
            }
            if (shared == -1) continue;
            angles.push_back({shared, a1, a2});
            const double cos_t = vdot(vnormalized(P[a1] - P[shared]), vnormalized(P[a2] - P[shared]));
            interactions_.emplace_back(new HarmonicAngleInteraction(shared, a1, a2, ip.angle_k, std::acos(cos_t)));
        }
    std::vector<char> improper(n, 0);   // improper centres already used
    for (size_t i = 0; i < angles.size(); ++i)
        for (size_t j = i + 1; j < angles.size(); ++j) {
            const std::array<int, 3>& A = angles[i];
            const std::array<int, 3>& B = angles[j];
            int a1 = -1, a2 = -1, a3 = -1, a4 = -1;
            if (A[0] == B[0] && !improper[A[0]]) {
                improper[A[0]] = 1;
                a1 = A[0]; a2 = A[1]; a3 = A[2];
                a4 = (B[1] == A[1] || B[1] == A[2]) ? B[2] : B[1];
            } else if (A[1] == B[0] && B[1] == A[0]) {
                a1 = A[2]; a2 = A[0]; a3 = B[0]; a4 = B[2];
            } else if (A[2] == B[0] && B[1] == A[0]) {
                a1 = A[1]; a2 = A[0]; a3 = B[0]; a4 = B[2];
            } else if (A[1] == B[0] && B[2] == A[0]) {
                a1 = A[2]; a2 = A[0]; a3 = B[0]; a4 = B[1];
            } else if (A[2] == B[0] && B[2] == A[0]) {
                a1 = A[1]; a2 = A[0]; a3 = B[0]; a4 = B[1];
            }
            if (a1 == -1) continue;
            const Vec3d d12 = P[a1] - P[a2], d23 = P[a2] - P[a3], d43 = P[a4] - P[a3];
            const Vec3d m = vcross(d12, d23), nv = vcross(d43, d23);
            const double tan_s = vdot(nv, d12) * vlen(d23) / vdot(m, nv);
            interactions_.emplace_back(
                new HarmonicDihedralInteraction(a1, a2, a3, a4, ip.dihedral_k, std::atan(tan_s)));
        }
    for (int i = 2; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            if (component[i] >= 0 && component[i] == component[j]) continue;
            addVDWInteraction(i, j);
            if (options_.coulomb) addCoulombInteraction(i, j);
        }
}

void Simulation::initialize() {                   // simulation.cpp:31-65
    if (options_.normal == NORMAL_X) system.rotateCoordAxes("YZX");
    else if (options_.normal == NORMAL_Y) system.rotateCoordAxes("ZXY");
    if (options_.vdw_pbc) {
        std::vector<Vec3d> cv;
        if (options_.use_external_potential) {
            Cube c = readCube(options_.e_potential_file, false);
            for (int a = 0; a < 3; ++a) cv.push_back((double)c.n[a] * c.voxel[a]);
        } else {
            cv = {options_.cell_a, options_.cell_b, options_.cell_c};
        }
        system.setUnitCell(cv);
    }
    system.centerMolecule(options_.center);
    system.setMoleculeZ();
    calculateTipDummyDistance();
    buildInteractions();
    if (options_.flexible) lowerFlexSystem();
    else lowerSystem();
    initialized_ = true;
}

void Simulation::lowerFlexSystem() {
    flex_terms_.clear();
    n_tip_surface_terms_ = 0;
    for (const auto& it : interactions_) {
        if (it->isTipSurface() && n_tip_surface_terms_ == (int)flex_terms_.size()) ++n_tip_surface_terms_;
        it->lowerFlex(flex_terms_);
    }
    const int n = system.n_atoms_;
    flex_pos_.assign(3 * (size_t)n, 0.0);
    flex_fixed_.assign(n, 0);
    for (int i = 0; i < n; ++i) {
        flex_pos_[3 * i] = system.positions_[i].x;
        flex_pos_[3 * i + 1] = system.positions_[i].y;
        flex_pos_[3 * i + 2] = system.positions_[i].z;
        flex_fixed_[i] = system.fixed_[i];
    }
    flex_desc_ = mafm_flex_desc{};
    flex_desc_.n_atoms = n;
    flex_desc_.n_terms = (int32_t)flex_terms_.size();
    flex_desc_.positions = flex_pos_.data();
    flex_desc_.fixed = flex_fixed_.data();
    flex_desc_.terms = flex_terms_.data();
    flex_desc_.tip_pbc = system.tipPbc() ? 1 : 0;
    // current_point == total_points / 2 (simulation.cpp:114)
    flex_desc_.state_column = (n_points_[0] * n_points_[1]) / 2;
    flex_desc_.tip_dummy_d = system.getTipDummyDistance();
    memcpy(flex_desc_.cell, system.getUnitCell().m, sizeof flex_desc_.cell);
    flex_desc_.offset[0] = system.getOffset().x;
    flex_desc_.offset[1] = system.getOffset().y;
    flex_desc_.offset[2] = system.getOffset().z;
}

void Simulation::lowerSystem() {
    low_ = Lowering();
    low_.positions = &system.positions_;
    for (const auto& it : interactions_) it->lower(low_);
    if (!low_.tip_dummy_vdw_set) fail(MAFM_ERR_STATE, "no tip-dummy VDW term in the interaction list");
    factorImages(low_);
    desc_ = mafm_system_desc{};
    desc_.n_sites = (int32_t)low_.sites.size();
    desc_.sites = low_.sites.data();
    desc_.n_images = low_.n_images;
    desc_.image_shifts = low_.image_shifts.empty() ? nullptr : low_.image_shifts.data();
    desc_.force_grid = low_.force_grid;
    desc_.potential_grid = low_.potential_grid;
    desc_.tip_pbc = (system.tipPbc() && !options_.rigidgrid) ? 1 : 0;   // setTipPbc(false), simulation.cpp:98-99
    desc_.tip_dummy = low_.tip_dummy;
    desc_.tip_dummy_d = system.getTipDummyDistance();
    memcpy(desc_.cell, system.getUnitCell().m, sizeof desc_.cell);
    desc_.offset[0] = system.getOffset().x;
    desc_.offset[1] = system.getOffset().y;
    desc_.offset[2] = system.getOffset().z;
}

mafm_scan_desc Simulation::scanDesc(const mafm_run_opts& r) const {
    const InputOptions& o = options_;
    mafm_scan_desc d{};
    d.nx = n_points_[0];
    d.ny = n_points_[1];
    d.nz = n_points_[2];
    d.dx = o.dx; d.dy = o.dy; d.dz = o.dz;
    d.zhigh = o.zhigh;
    d.minimiser = o.minimiser_type == FIRE ? MAFM_MIN_FIRE : MAFM_MIN_SD;
    d.integrator = o.integrator_type == EULER ? MAFM_INT_EULER : (o.integrator_type == RK4 ? MAFM_INT_RK4 : MAFM_INT_MIDPOINT);
    d.minterm = o.minterm == MIN_E ? MAFM_TERM_E : (o.minterm == MIN_F ? MAFM_TERM_F : MAFM_TERM_EF);
    d.etol = o.etol; d.ftol = o.ftol; d.dt = o.dt;
    d.maxsteps = o.maxsteps;
    d.precision = r.precision;
    d.lanes_per_pixel = r.lanes_per_pixel;
    d.waves_per_simd = r.waves_per_simd;
    d.compensation = r.compensation;
    d.waves_per_group = r.waves_per_group;
    d.options = r.options;
    return d;
}

void Simulation::run(const mafm_run_opts& r, mafm_scan_stats* stats) {   // simulation.cpp:67-143
    if (!initialized_) fail(MAFM_ERR_STATE, "run before initialize");
    mafm_scan_desc d = scanDesc(r);
    scan_generation_ = 0;
    if (options_.flexible) {
        check(ctx_, mafm_scan_flexible(ctx_, &flex_desc_, &d, &last_stats_), "flexible scan");
    } else {
        check(ctx_, mafm_set_systems(ctx_, 1, &desc_), "set system");
        check(ctx_, mafm_scan(ctx_, &d, &last_stats_), "scan");
    }
    record_base_ = 0;
    system_index_ = 0;
    scan_generation_ = ctx_scan_generation(ctx_);
    if (stats) *stats = last_stats_;
}

void Simulation::requireOwnRecords() const {
    if (host_records_)
        fail(MAFM_ERR_STATE, "the records of a streamed batch run are on the host only; device-side "
                             "post-processing needs the job run again");
    if (scan_generation_ == 0) fail(MAFM_ERR_STATE, "no records: this simulation has not run");
    if (scan_generation_ != ctx_scan_generation(ctx_))
        fail(MAFM_ERR_STATE, "the context has run another scan since this simulation's: its records are gone "
                             "(run the simulation again)");
}

void Simulation::fetchRecords() {
    if (host_records_) return;   // a streamed batch left them on the host
    requireOwnRecords();
    size_t n = (size_t)n_points_[0] * n_points_[1] * n_points_[2];
    records_.resize(n);
    check(ctx_, mafm_records(ctx_, record_base_, n, records_.data()), "records");
}

// openUniverse / writeOutput / closeUniverse (mechafm.cpp:59-101,
// simulation.cpp:145-199): one text file per height, fixed printf formats,
// records in (i, j) order; each file is formatted on its own thread.
// gzip stream of `data` into `path` (deflate level 6, the level of the
// reference's "gzip -6" pipe, mechafm.cpp:63-66); false on an I/O error.
static bool write_gzip(const std::string& path, const std::string& data) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return false;
    z_stream zs{};
    if (deflateInit2(&zs, 6, Z_DEFLATED, 15 + 16, 8, Z_DEFAULT_STRATEGY) != Z_OK) {
        fclose(f);
        return false;
    }
    std::vector<unsigned char> out(1 << 20);
    zs.next_in = (Bytef*)data.data();
    zs.avail_in = (uInt)data.size();
    bool ok = true;
    int rc;
    do {
        zs.next_out = out.data();
        zs.avail_out = (uInt)out.size();
        rc = deflate(&zs, Z_FINISH);
        const size_t have = out.size() - zs.avail_out;
        if (have && fwrite(out.data(), 1, have, f) != have) ok = false;
    } while (ok && rc == Z_OK);
    deflateEnd(&zs);
    if (rc != Z_STREAM_END) ok = false;
    if (fclose(f) != 0) ok = false;
    return ok;
}

void Simulation::writeOutput(int threads) {
    if (records_.empty()) fetchRecords();
    const int nz = n_points_[2];
    std::vector<std::string> errors(nz);
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = std::max(1, std::min(threads, nz));
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&]() {
            for (int k = next++; k < nz; k = next++) errors[k] = writeHeight(k, records_.data());
        });
    for (auto& th : pool) th.join();
    for (const auto& e : errors)
        if (!e.empty()) fail(MAFM_ERR_INVALID, "%s", e.c_str());
    writeStatistics(records_.data());
}

// The lines of height k for pixel columns [c0, c1) (column i*ny + j),
// appended to `out` as the reference prints them (simulation.cpp:180-195):
// "%d %d %d %6.3f %6.3f %6.3f %8.4f %8.4f %8.4f %6.3f %6.3f %6.3f %6.3f %8.4f
// %8.4f %d"
void Simulation::formatLines(int k, const mafm_record* recs, size_t c0, size_t c1, std::string& out) const {
    const int ny = n_points_[1], nz = n_points_[2];
    size_t used = out.size();
    out.resize(used + (c1 - c0) * 160 + 4096);
    char* p = &out[0] + used;
    for (size_t col = c0; col < c1; ++col) {
        const int i = (int)(col / ny), j = (int)(col % ny);
        const mafm_record& r = recs[col * nz + k];
        const double rl = std::sqrt(std::pow(r.rx, 2) + std::pow(r.ry, 2) + std::pow(r.rz, 2));
        const double ang = std::atan2(std::sqrt(std::pow(r.rx, 2) + std::pow(r.ry, 2)), r.rz) * (180.0 / PI);
        p = fmt_int(p, k); *p++ = ' ';
        p = fmt_int(p, i); *p++ = ' ';
        p = fmt_int(p, j); *p++ = ' ';
        p = fmt_fixed<3>(p, r.x, 6); *p++ = ' ';
        p = fmt_fixed<3>(p, r.y, 6); *p++ = ' ';
        p = fmt_fixed<3>(p, r.z, 6); *p++ = ' ';
        p = fmt_fixed<4>(p, r.fx, 8); *p++ = ' ';
        p = fmt_fixed<4>(p, r.fy, 8); *p++ = ' ';
        p = fmt_fixed<4>(p, r.fz, 8); *p++ = ' ';
        p = fmt_fixed<3>(p, r.rx, 6); *p++ = ' ';
        p = fmt_fixed<3>(p, r.ry, 6); *p++ = ' ';
        p = fmt_fixed<3>(p, r.rz, 6); *p++ = ' ';
        p = fmt_fixed<3>(p, rl, 6); *p++ = ' ';
        p = fmt_fixed<4>(p, ang, 8); *p++ = ' ';
        p = fmt_fixed<4>(p, r.e, 8); *p++ = ' ';
        p = fmt_int(p, r.n_steps); *p++ = '\n';
        if ((size_t)(p - &out[0]) + 4096 > out.size()) {   // non-finite or huge values: grow
            used = (size_t)(p - &out[0]);
            out.resize(out.size() * 2);
            p = &out[0] + used;
        }
    }
    out.resize((size_t)(p - &out[0]));
}

std::string Simulation::heightPath(int k) const {
    char name[512];
    snprintf(name, sizeof name, "%sscan-%06.3f.dat", options_.outputfolder.c_str(), options_.zhigh - k * options_.dz);
    return options_.gzip ? std::string(name) + ".gz" : std::string(name);
}

// One file per height, lines in (i, j) order.
std::string Simulation::writeHeight(int k, const mafm_record* recs) const {
    std::string body;
    formatLines(k, recs, 0, (size_t)n_points_[0] * n_points_[1], body);
    const std::string name = heightPath(k);
    if (options_.gzip) return write_gzip(name, body) ? std::string() : "cannot write " + name;
    FILE* f = fopen(name.c_str(), "wb");
    if (!f) return "cannot open output " + name;
    std::string err;
    if (fwrite(body.data(), 1, body.size(), f) != body.size()) err = "short write to " + name;
    if (fclose(f) != 0 && err.empty()) err = "cannot close " + name;
    return err;
}

void Simulation::writeStatistics(const mafm_record* recs) const {
    const InputOptions& o = options_;
    if (!o.statistics) return;                    // finalize (mechafm.cpp:129-152)
    const int64_t n_pts = (int64_t)n_points_[0] * n_points_[1] * n_points_[2];
    std::string path = o.outputfolder + "statistics.txt";
    FILE* f = fopen(path.c_str(), "w");
    if (!f) return;
    fprintf(f, "Simulation run finished\n");
    fprintf(f, "Statistics:\n");
    fprintf(f, "    Computed %lld tip positions\n", (long long)n_pts);
    // this job's own iterations (a batched launch's statistics cover all its jobs)
    long long steps = 0;
    for (int64_t q = 0; q < n_pts; ++q) steps += recs[q].n_steps;
    fprintf(f, "    Needed %lld minimization steps in total\n", steps);
    fprintf(f, "    Which means approximately %.2f minimization steps per tip position\n", (double)steps / n_pts);
    // the reference's two times (mechafm.cpp:137-150), from the job's opening
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - opened_).count();
    fprintf(f, "    The simulation wall time is %.2f seconds\n", wall);
    fprintf(f, "    The entire simulation took %.2f seconds\n", wall);
    fclose(f);
}

// ---------------------------------------------------------------- batch writer
// Writes the per-height files of many jobs on one pool of threads, a chunk of
// pixel columns at a time.  The lines of one chunk at one height are
// formatted as soon as the chunk's records of that height are on the host
// (ready()); with
// gzip on, each such segment is deflated on its own as raw deflate ending on
// a byte boundary (Z_SYNC_FLUSH; the file's last segment Z_FINISH), so the
// segments of a file concatenate into one deflate stream -- the layout of
// pigz -- behind one gzip header, with the CRC-32 combined from the
// segments'.  A file is written by whichever thread completes its last
// segment.
class BatchWriter {
  public:
    // recs[m]: job m's records in mafm_records' order; copy: also copy them
    // into the job's host records once the whole job is in
    BatchWriter(std::vector<Simulation*> sims, std::vector<const mafm_record*> recs, int threads, bool copy)
        : sims_(std::move(sims)), recs_(std::move(recs)), copy_(copy) {
        const int n = (int)sims_.size();
        jobs_.resize(n);
        for (int m = 0; m < n; ++m) {
            Job& j = jobs_[m];
            const Simulation& s = *sims_[m];
            j.cols = (size_t)s.n_points_[0] * s.n_points_[1];
            j.nz = s.n_points_[2];
            j.chunk = std::max<size_t>(256, (j.cols + 63) / 64);
            j.n_chunks = (int)((j.cols + j.chunk - 1) / j.chunk);
            j.queued.assign(j.n_chunks, 0);
            j.files.reset(new File[j.nz]);
            for (int k = 0; k < j.nz; ++k) {
                j.files[k].seg.resize(j.n_chunks);
                j.files[k].crc.resize(j.n_chunks);
                j.files[k].len.resize(j.n_chunks);
            }
        }
        if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
        threads = std::max(1, std::min(threads, 64));
        for (int t = 0; t < threads; ++t) pool_.emplace_back([this]() { work(); });
    }
    ~BatchWriter() { close(); }

    int chunks(int m) const { return jobs_[m].n_chunks; }
    // pixel columns [first, last) of chunk c of job m
    size_t chunk_first(int m, int c) const { return (size_t)c * jobs_[m].chunk; }
    size_t chunk_last(int m, int c) const { return std::min(jobs_[m].cols, (size_t)(c + 1) * jobs_[m].chunk); }
    // heights of chunk c of job m handed to the writers so far
    int queued(int m, int c) const { return jobs_[m].queued[c]; }
    // every height of every chunk of job m has been handed over
    bool complete(int m) const { return jobs_[m].full == jobs_[m].n_chunks; }

    // the records of heights [0, heights) of every column of chunk c of job
    // m are on the host (called from one thread)
    void ready(int m, int c, int heights) {
        Job& j = jobs_[m];
        heights = std::min(heights, j.nz);
        if (heights <= j.queued[c]) return;
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (int k = j.queued[c]; k < heights; ++k) tasks_.push_back({m, c, k});
            if (heights == j.nz && ++j.full == j.n_chunks) tasks_.push_back({m, -1, -1});   // copy, statistics
        }
        j.queued[c] = heights;
        cv_.notify_all();
    }

    // waits for every queued task; throws the first error
    void finish() {
        close();
        if (!err_.empty()) throw Error(MAFM_ERR_INVALID, err_);
    }

  private:
    struct File {
        std::vector<std::string> seg;
        std::vector<uint32_t> crc;
        std::vector<size_t> len;
        std::atomic<int> done{0};
    };
    struct Job {
        size_t cols = 0, chunk = 1;
        int nz = 0, n_chunks = 0, full = 0;   // full: chunks with every height handed over
        std::vector<int> queued;
        std::unique_ptr<File[]> files;
    };
    struct Task {
        int m, c, k;
    };

    void close() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            closing_ = true;
        }
        cv_.notify_all();
        for (auto& t : pool_) t.join();
        pool_.clear();
    }

    void error(const std::string& e) {
        std::lock_guard<std::mutex> lk(mu_);
        if (err_.empty()) err_ = e;
    }

    void work() {
        for (;;) {
            Task t;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&]() { return closing_ || !tasks_.empty(); });
                if (tasks_.empty()) return;
                t = tasks_.front();
                tasks_.pop_front();
            }
            try {
                run(t);
            } catch (const std::exception& e) {
                error(e.what());
            }
        }
    }

    void run(const Task& t) {
        Job& j = jobs_[t.m];
        Simulation& sim = *sims_[t.m];
        const mafm_record* r = recs_[t.m];
        if (t.c < 0) {
            if (copy_) {
                sim.records_.assign(r, r + j.cols * j.nz);
                sim.host_records_ = true;
            }
            sim.writeStatistics(r);
            return;
        }
        File& f = j.files[t.k];
        const size_t c0 = (size_t)t.c * j.chunk, c1 = std::min(j.cols, c0 + j.chunk);
        std::string text;
        sim.formatLines(t.k, r, c0, c1, text);
        f.len[t.c] = text.size();
        if (sim.options_.gzip) {
            f.crc[t.c] = (uint32_t)crc32(0L, (const Bytef*)text.data(), (uInt)text.size());
            f.seg[t.c] = deflate_segment(text, t.c + 1 == j.n_chunks);
        } else {
            f.seg[t.c] = std::move(text);
        }
        if (f.done.fetch_add(1) + 1 == j.n_chunks) write_file(sim, t.k, f);
    }

    static std::string deflate_segment(const std::string& in, bool last) {
        z_stream zs{};
        if (deflateInit2(&zs, 6, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK)
            throw std::runtime_error("deflateInit2 failed");
        std::string out;
        out.resize(deflateBound(&zs, (uLong)in.size()) + 16);
        zs.next_in = (Bytef*)in.data();
        zs.avail_in = (uInt)in.size();
        zs.next_out = (Bytef*)&out[0];
        zs.avail_out = (uInt)out.size();
        const int rc = deflate(&zs, last ? Z_FINISH : Z_SYNC_FLUSH);
        const size_t have = out.size() - zs.avail_out;
        deflateEnd(&zs);
        if ((last && rc != Z_STREAM_END) || (!last && (rc != Z_OK || zs.avail_in != 0)))
            throw std::runtime_error("deflate failed");
        out.resize(have);
        return out;
    }

    void write_file(const Simulation& sim, int k, File& f) {
        const std::string name = sim.heightPath(k);
        FILE* fp = fopen(name.c_str(), "wb");
        if (!fp) {
            error("cannot open output " + name);
            return;
        }
        bool ok = true;
        const bool gz = sim.options_.gzip;
        if (gz) {   // gzip header: deflate, no name, mtime 0, Unix (what zlib's gzip wrapper writes)
            static const unsigned char hdr[10] = {0x1f, 0x8b, 8, 0, 0, 0, 0, 0, 0, 3};
            ok = fwrite(hdr, 1, sizeof hdr, fp) == sizeof hdr;
        }
        uLong crc = 0;
        uint64_t total = 0;
        for (size_t c = 0; c < f.seg.size() && ok; ++c) {
            ok = fwrite(f.seg[c].data(), 1, f.seg[c].size(), fp) == f.seg[c].size();
            crc = c == 0 ? f.crc[0] : crc32_combine(crc, f.crc[c], (z_off_t)f.len[c]);
            total += f.len[c];
            std::string().swap(f.seg[c]);
        }
        if (gz && ok) {   // trailer: CRC-32 and length mod 2^32, little-endian
            unsigned char tr[8];
            for (int b = 0; b < 4; ++b) {
                tr[b] = (unsigned char)(crc >> (8 * b));
                tr[4 + b] = (unsigned char)(total >> (8 * b));
            }
            ok = fwrite(tr, 1, 8, fp) == 8;
        }
        if (fclose(fp) != 0) ok = false;
        if (!ok) error("cannot write " + name);
    }

    std::vector<Simulation*> sims_;
    std::vector<const mafm_record*> recs_;
    bool copy_;
    std::vector<Job> jobs_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<Task> tasks_;
    bool closing_ = false;
    std::string err_;
    std::vector<std::thread> pool_;
};

void Simulation::writeStates() {                  // system.cpp:54-65
    if (!options_.flexible || flex_desc_.state_column < 0) return;
    requireOwnRecords();
    if (records_.empty()) fetchRecords();
    const int n = system.n_atoms_, nz = n_points_[2];
    std::vector<double> st((size_t)nz * n * 3);
    check(ctx_, mafm_flex_state(ctx_, 0, nz, st.data()), "flexible state");
    for (int k = 0; k < nz; ++k) {
        const mafm_record& r = records_[(size_t)flex_desc_.state_column * nz + k];
        const double* q = &st[(size_t)k * n * 3];
        char name[1024];
        snprintf(name, sizeof name, "%sstate_%.1f-%.1f-%.1f.xyz", options_.outputfolder.c_str(), r.x, r.y, q[2]);
        std::string body = std::to_string(n) + "\n\n";
        char line[256];
        for (int i = 0; i < n; ++i) {
            snprintf(line, sizeof line, "%s %8.4f %8.4f %8.4f\n", system.types_[i].c_str(), q[3 * i], q[3 * i + 1],
                     q[3 * i + 2]);
            body += line;
        }
        FILE* f = fopen(name, "w");
        if (!f) fail(MAFM_ERR_INVALID, "cannot write %s", name);
        const bool ok = fwrite(body.data(), 1, body.size(), f) == body.size();
        if (fclose(f) != 0 || !ok) fail(MAFM_ERR_INVALID, "cannot write %s", name);
        log_ += "+- Wrote ";
        log_ += name;
        log_ += "\n";
    }
}

void Simulation::writeDeltaF(const mafm_deltaf_desc& d) {
    const InputOptions& o = options_;
    const int nx = n_points_[0], ny = n_points_[1], nz = n_points_[2];
    mafm_deltaf_weights_out w;
    check(ctx_, mafm_deltaf_weights(ctx_, &d, nz, o.zhigh, o.dz, &w), "frequency-shift window");
    if (w.n_images == 0) fail(MAFM_ERR_INVALID, "%d heights do not fill a %d-height window", nz, w.n_weights);
    std::vector<double> df((size_t)w.n_images * nx * ny), zlow(w.n_images);
    requireOwnRecords();
    check(ctx_, mafm_delta_f(ctx_, system_index_, &d, df.data(), zlow.data()), "frequency shift");
    for (int p = 0; p < w.n_images; ++p) {
        char name[512];
        snprintf(name, sizeof name, "%sdeltaf-%06.3f.dat", o.outputfolder.c_str(), zlow[p]);
        FILE* f = fopen(name, "w");
        if (!f) fail(MAFM_ERR_INVALID, "cannot open output %s", name);
        fprintf(f, "%% i j x y delta_f[Hz] (A = %.3f A, window of %d heights, dz = %.4f A)\n", w.amplitude,
                w.n_weights, w.dz);
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                fprintf(f, "%d %d %6.3f %6.3f %.6e\n", i, j, i * o.dx, j * o.dy, df[((size_t)p * nx + i) * ny + j]);
        fclose(f);
    }
}

}  // namespace mafm

// ================================================================ C ABI
using mafm::Error;
using mafm::Simulation;

// A simulation's reference on its context (released after the simulation's
// own destructor has freed the grids it built there).
struct CtxRef {
    mafm_ctx* c;
    explicit CtxRef(mafm_ctx* x) : c(x) { mafm::ctx_retain(c); }
    ~CtxRef() { mafm::ctx_release(c); }
    CtxRef(const CtxRef&) = delete;
    CtxRef& operator=(const CtxRef&) = delete;
};

struct mafm_sim {
    CtxRef ref;   // declared first, destroyed last
    Simulation sim;
    std::string err;
    explicit mafm_sim(mafm_ctx* c) : ref(c), sim(c) {}
};

template <typename F>
static mafm_status guarded(mafm_sim* s, F&& f) {
    try {
        f();
        return MAFM_OK;
    } catch (const Error& e) {
        if (s) s->err = e.what();
        return e.status;
    } catch (const std::exception& e) {
        if (s) s->err = e.what();
        return MAFM_ERR_INVALID;
    }
}

extern "C" {

mafm_status mafm_sim_open(const char* input, const char* outdir, mafm_ctx* ctx, mafm_sim** out) {
    if (!input || !out) return MAFM_ERR_INVALID;
    *out = nullptr;
    mafm_sim* s = new mafm_sim(ctx);
    mafm_status st = guarded(s, [&]() {
        s->sim.parseCommandLine(input, outdir ? outdir : "");
        s->sim.readInputFile();
        s->sim.readXYZFile();
        s->sim.readParameterFile();
    });
    *out = s;   // returned even on failure so the caller can read the message
    return st;
}

mafm_status mafm_sim_initialize(mafm_sim* s) {
    if (!s) return MAFM_ERR_INVALID;
    return guarded(s, [&]() { s->sim.initialize(); });
}

mafm_status mafm_sim_info_get(mafm_sim* s, mafm_sim_info* info) {
    if (!s || !info) return MAFM_ERR_INVALID;
    memset(info, 0, sizeof *info);
    const Simulation& m = s->sim;
    info->nx = m.n_points_[0];
    info->ny = m.n_points_[1];
    info->nz = m.n_points_[2];
    info->n_atoms = m.system.n_atoms_ - 2;
    info->n_sites = m.options_.flexible ? m.n_tip_surface_terms_ : m.systemDesc().n_sites;
    info->flexible = m.options_.flexible;
    info->n_terms = m.options_.flexible ? m.flexDesc().n_terms : 0;
    info->rigidgrid = m.options_.rigidgrid;
    info->use_external_potential = m.options_.use_external_potential;
    info->vdw_pbc = m.options_.vdw_pbc;
    info->coulomb = m.options_.coulomb;
    for (int a = 0; a < 3; ++a) info->force_grid_n[a] = m.force_grid_n_[a];
    info->tip_dummy_d = m.system.getTipDummyDistance();
    info->offset[0] = m.system.getOffset().x;
    info->offset[1] = m.system.getOffset().y;
    info->offset[2] = m.system.getOffset().z;
    info->grid_build_ms = m.grid_build_ms_;
    return MAFM_OK;
}

mafm_status mafm_sim_system(mafm_sim* s, mafm_system_desc* d) {
    if (!s || !d) return MAFM_ERR_INVALID;
    if (!s->sim.initialized_) {
        s->err = "system requested before initialize";
        return MAFM_ERR_STATE;
    }
    if (s->sim.options_.flexible) {
        s->err = "a flexible job has a term list (mafm_sim_flex_system), not a site table";
        return MAFM_ERR_STATE;
    }
    *d = s->sim.systemDesc();
    return MAFM_OK;
}

mafm_status mafm_sim_flex_system(mafm_sim* s, mafm_flex_desc* d) {
    if (!s || !d) return MAFM_ERR_INVALID;
    if (!s->sim.initialized_) {
        s->err = "system requested before initialize";
        return MAFM_ERR_STATE;
    }
    if (!s->sim.options_.flexible) {
        s->err = "a rigid job has a site table (mafm_sim_system), not a term list";
        return MAFM_ERR_STATE;
    }
    *d = s->sim.flexDesc();
    return MAFM_OK;
}

mafm_status mafm_sim_scan_desc(mafm_sim* s, const mafm_run_opts* o, mafm_scan_desc* d) {
    if (!s || !d) return MAFM_ERR_INVALID;
    mafm_run_opts def{};
    def.precision = MAFM_PREC_FAST64;
    *d = s->sim.scanDesc(o ? *o : def);
    return MAFM_OK;
}

mafm_status mafm_sim_run(mafm_sim* s, const mafm_run_opts* o, mafm_scan_stats* stats) {
    if (!s) return MAFM_ERR_INVALID;
    mafm_run_opts def{};
    def.precision = MAFM_PREC_FAST64;
    return guarded(s, [&]() {
        s->sim.records_.clear();
        s->sim.host_records_ = false;
        s->sim.run(o ? *o : def, stats);
    });
}

// The members of a batched launch: initialised rigid jobs on one context
// with the same scan geometry and minimiser settings; returns the shared
// scan descriptor.
static mafm_scan_desc batch_desc(int n, mafm_sim* const* sims, const mafm_run_opts& o) {
    mafm_ctx* ctx = sims[0]->sim.ctx_;
    if (!ctx) throw Error(MAFM_ERR_STATE, "a batch needs a device context");
    mafm_scan_desc d0 = sims[0]->sim.scanDesc(o);
    for (int m = 0; m < n; ++m) {
        if (!sims[m]) throw Error(MAFM_ERR_INVALID, "batch member is NULL");
        Simulation& s = sims[m]->sim;
        if (!s.initialized_) throw Error(MAFM_ERR_STATE, "batch member not initialised");
        if (s.options_.flexible) throw Error(MAFM_ERR_INVALID, "flexible jobs run one at a time (mafm_sim_run)");
        if (s.ctx_ != ctx) throw Error(MAFM_ERR_INVALID, "batch members must share one context");
        mafm_scan_desc d = s.scanDesc(o);
        if (d.nx != d0.nx || d.ny != d0.ny || d.nz != d0.nz || d.dx != d0.dx || d.dy != d0.dy ||
            d.dz != d0.dz || d.zhigh != d0.zhigh || d.minimiser != d0.minimiser ||
            d.integrator != d0.integrator || d.minterm != d0.minterm || d.etol != d0.etol ||
            d.ftol != d0.ftol || d.dt != d0.dt || d.maxsteps != d0.maxsteps)
            throw Error(MAFM_ERR_INVALID, "batch members must share the scan options");
    }
    return d0;
}

// One batched launch of the given members; their records stay on the device.
static void batch_scan(int n, mafm_sim* const* sims, const mafm_scan_desc& d0, mafm_scan_stats& ss) {
    mafm_ctx* ctx = sims[0]->sim.ctx_;
    std::vector<mafm_system_desc> descs;
    for (int m = 0; m < n; ++m) {
        descs.push_back(sims[m]->sim.systemDesc());
        sims[m]->sim.records_.clear();
        sims[m]->sim.host_records_ = false;
        sims[m]->sim.scan_generation_ = 0;
    }
    mafm_status st = mafm_set_systems(ctx, n, descs.data());
    if (st != MAFM_OK) throw Error(st, mafm_last_error(ctx));
    ss = mafm_scan_stats{};
    st = mafm_scan(ctx, &d0, &ss);
    if (st != MAFM_OK) throw Error(st, mafm_last_error(ctx));
    const size_t per = (size_t)d0.nx * d0.ny * d0.nz;
    for (int m = 0; m < n; ++m) {
        sims[m]->sim.record_base_ = per * m;
        sims[m]->sim.system_index_ = m;
        sims[m]->sim.last_stats_ = ss;
        sims[m]->sim.scan_generation_ = mafm::ctx_scan_generation(ctx);
    }
}

mafm_status mafm_batch_run(int n, mafm_sim* const* sims, const mafm_run_opts* o, mafm_scan_stats* stats) {
    if (n <= 0 || !sims || !sims[0]) return MAFM_ERR_INVALID;
    mafm_run_opts def{};
    def.precision = MAFM_PREC_FAST32;
    return guarded(sims[0], [&]() {
        const mafm_scan_desc d0 = batch_desc(n, sims, o ? *o : def);
        mafm_scan_stats ss{};
        batch_scan(n, sims, d0, ss);
        if (stats) *stats = ss;
    });
}

// One streamed launch of n jobs with their files written as the records
// land; progress (if any) gets the fraction of positions handed to the
// writers, from the polling thread, then 1 at the end.
static mafm_status run_output(int n, mafm_sim* const* sims, const mafm_run_opts* o, int threads,
                              mafm_progress_fn progress, void* user, mafm_scan_stats* stats) {
    if (n <= 0 || !sims || !sims[0]) return MAFM_ERR_INVALID;
    mafm_run_opts def{};
    def.precision = MAFM_PREC_FAST64;
    return guarded(sims[0], [&]() {
        const mafm_scan_desc d0 = batch_desc(n, sims, o ? *o : def);
        mafm_ctx* ctx = sims[0]->sim.ctx_;
        std::vector<mafm_system_desc> descs;
        for (int m = 0; m < n; ++m) {
            descs.push_back(sims[m]->sim.systemDesc());
            sims[m]->sim.records_.clear();
            sims[m]->sim.host_records_ = false;
            sims[m]->sim.scan_generation_ = 0;
        }
        mafm_status st = mafm_set_systems(ctx, n, descs.data());
        if (st != MAFM_OK) throw Error(st, mafm_last_error(ctx));
        const mafm_record* recs = nullptr;
        const int32_t* done = nullptr;
        st = mafm_scan_host(ctx, &d0, &recs, &done);
        if (st != MAFM_OK) throw Error(st, mafm_last_error(ctx));
        struct EndScan {   // the launch is ended on every way out
            mafm_ctx* c;
            bool armed = true;
            ~EndScan() {
                mafm_scan_stats ignored;
                if (armed) (void)mafm_scan_host_wait(c, &ignored);
            }
        } end_scan{ctx};
        const size_t cols = (size_t)d0.nx * d0.ny, per = cols * d0.nz;

        // the writers format each chunk of columns, height by height, as soon
        // as it is in host memory, straight from the pinned records, while the
        // rest scans
        std::vector<Simulation*> jobs;
        std::vector<const mafm_record*> rp;
        for (int m = 0; m < n; ++m) {
            jobs.push_back(&sims[m]->sim);
            rp.push_back(recs + per * m);
        }
        mafm::BatchWriter writer(jobs, rp, threads, true);
        // poll the column counters (acquire: a column's records of heights
        // [0, count) are visible once the count is); each chunk of columns
        // goes to the writers height by height, as its slowest column passes
        const int nz = d0.nz;
        const double total_cells = (double)n * cols * nz;
        double handed = 0;
        std::vector<char> job_done(n, 0);
        int complete = 0;
        bool finished = false;
        mafm_status sst = MAFM_OK;
        std::string serr;
        while (complete < n) {
            bool moved = false;
            for (int m = 0; m < n; ++m) {
                if (job_done[m]) continue;
                const int32_t* f = done + cols * m;
                for (int c = 0; c < writer.chunks(m); ++c) {
                    const int q = writer.queued(m, c);
                    if (q == nz) continue;
                    int h = nz;
                    for (size_t col = writer.chunk_first(m, c), e = writer.chunk_last(m, c); col < e && h > q; ++col)
                        h = std::min(h, __atomic_load_n(f + col, __ATOMIC_ACQUIRE));
                    if (h > q) {
                        writer.ready(m, c, h);
                        handed += (double)(h - q) * (writer.chunk_last(m, c) - writer.chunk_first(m, c));
                        moved = true;
                    }
                }
                if (writer.complete(m)) {
                    job_done[m] = 1;
                    ++complete;
                }
            }
            if (moved && progress) progress(handed / total_cells, user);
            if (complete == n || moved) continue;
            if (finished) {
                sst = MAFM_ERR_HIP;
                serr = "the streamed scan finished with columns unmarked";
                break;
            }
            int fin = 0;
            st = mafm_scan_host_query(ctx, &fin);
            if (st != MAFM_OK) {
                sst = st;
                serr = mafm_last_error(ctx);
                break;
            }
            if (fin) finished = true;   // one more pass reads the last counters
            else std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
        mafm_scan_stats ss{};
        end_scan.armed = false;
        const mafm_status est = mafm_scan_host_wait(ctx, &ss);
        const std::string eerr = est != MAFM_OK ? mafm_last_error(ctx) : "";
        std::string werr;
        try {
            writer.finish();
        } catch (const Error& e) {
            werr = e.what();
        }
        if (sst != MAFM_OK) throw Error(sst, serr);
        if (est != MAFM_OK) throw Error(est, eerr);
        if (!werr.empty()) throw Error(MAFM_ERR_INVALID, werr);
        for (int m = 0; m < n; ++m) {
            sims[m]->sim.record_base_ = per * m;
            sims[m]->sim.system_index_ = m;
            sims[m]->sim.last_stats_ = ss;
        }
        if (progress) progress(1.0, user);
        if (stats) *stats = ss;
    });
}

mafm_status mafm_batch_run_output(int n, mafm_sim* const* sims, const mafm_run_opts* o, int threads,
                                  mafm_scan_stats* stats) {
    return run_output(n, sims, o, threads, nullptr, nullptr, stats);
}

mafm_status mafm_sim_run_output(mafm_sim* sim, const mafm_run_opts* o, int threads, mafm_progress_fn progress,
                                void* user, mafm_scan_stats* stats) {
    if (!sim) return MAFM_ERR_INVALID;
    mafm_sim* one[1] = {sim};
    return run_output(1, one, o, threads, progress, user, stats);
}

mafm_status mafm_batch_write_output(int n, mafm_sim* const* sims, int threads) {
    if (n <= 0 || !sims || !sims[0]) return MAFM_ERR_INVALID;
    return guarded(sims[0], [&]() {
        std::vector<Simulation*> jobs;
        std::vector<const mafm_record*> rp;
        for (int m = 0; m < n; ++m) {
            if (!sims[m]) throw Error(MAFM_ERR_INVALID, "batch member is NULL");
            Simulation& s = sims[m]->sim;
            if (s.options_.flexible) throw Error(MAFM_ERR_INVALID, "flexible jobs write with mafm_sim_write_output");
            if (s.records_.empty()) s.fetchRecords();
            jobs.push_back(&s);
            rp.push_back(s.records_.data());
        }
        mafm::BatchWriter writer(jobs, rp, threads, false);
        for (int m = 0; m < n; ++m)
            for (int c = 0; c < writer.chunks(m); ++c) writer.ready(m, c, jobs[m]->n_points_[2]);
        writer.finish();
    });
}

mafm_status mafm_sim_write_deltaf(mafm_sim* s, const mafm_deltaf_desc* d) {
    if (!s) return MAFM_ERR_INVALID;
    mafm_deltaf_desc def{};
    def.fz_from_text = 1;   // the script reads Fz back from the per-height files
    return guarded(s, [&]() { s->sim.writeDeltaF(d ? *d : def); });
}

mafm_status mafm_sim_settings(mafm_sim* s, char* buf, size_t size, size_t* needed) {
    if (!s || (size && !buf)) return MAFM_ERR_INVALID;
    const std::string t = s->sim.settingsText();
    if (needed) *needed = t.size() + 1;
    if (size) {
        const size_t n = std::min(size - 1, t.size());
        memcpy(buf, t.data(), n);
        buf[n] = 0;
    }
    return MAFM_OK;
}

mafm_status mafm_sim_records(mafm_sim* s, size_t first, size_t count, mafm_record* out) {
    if (!s) return MAFM_ERR_INVALID;
    return guarded(s, [&]() {
        if (s->sim.host_records_) {
            if (first + count > s->sim.records_.size() || (count && !out))
                throw Error(MAFM_ERR_INVALID, "records outside the job's");
            if (count) memcpy(out, s->sim.records_.data() + first, count * sizeof(mafm_record));
            return;
        }
        s->sim.requireOwnRecords();
        mafm_status st = mafm_records(s->sim.ctx_, s->sim.record_base_ + first, count, out);
        if (st != MAFM_OK) throw Error(st, mafm_last_error(s->sim.ctx_));
    });
}

mafm_status mafm_sim_write_output(mafm_sim* s, int threads) {
    if (!s) return MAFM_ERR_INVALID;
    return guarded(s, [&]() {
        s->sim.log_.clear();
        s->sim.fetchRecords();
        s->sim.writeStates();
        s->sim.writeOutput(threads);
    });
}

mafm_status mafm_sim_log(mafm_sim* s, char* buf, size_t size, size_t* needed) {
    if (!s || (size && !buf)) return MAFM_ERR_INVALID;
    const std::string& t = s->sim.log_;
    if (needed) *needed = t.size() + 1;
    if (size) {
        const size_t n = std::min(size - 1, t.size());
        memcpy(buf, t.data(), n);
        buf[n] = 0;
    }
    return MAFM_OK;
}

const char* mafm_sim_last_error(const mafm_sim* s) { return s ? s->err.c_str() : "no simulation"; }

void mafm_sim_close(mafm_sim* s) { delete s; }

}  // extern "C"
