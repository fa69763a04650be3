// C++ host API of the MI355X MechAFM scan: the reference's Simulation /
// System / Interaction / parser types (simulation.hpp, system.hpp,
// interactions.hpp, parse.hpp), rebuilt around the device tables of mafm.h.
//
// What stays as in the reference: option names and defaults, file formats,
// the set-up sequence (centre, lowest heavy atom at z=0, tip-dummy distance),
// the interaction list and its order, the output files.  What changes:
// Interaction objects are descriptors that lower themselves into device site
// tables instead of evaluating on the CPU, and Simulation::run hands the whole
// pixel / z / minimiser loop to one device launch.  Errors are thrown as
// mafm::Error and turned into mafm_status at the C ABI (no exit()).
#pragma once

#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <array>
#include <chrono>
#include <string>
#include <vector>

#include "../../../include/mafm.h"
#include "../../../include/mafm_sim.h"

namespace mafm {

struct Error : std::runtime_error {
    mafm_status status;
    Error(mafm_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

struct Vec3d {
    double x = 0, y = 0, z = 0;
    Vec3d() = default;
    Vec3d(double a, double b, double c) : x(a), y(b), z(c) {}
    Vec3d operator+(const Vec3d& o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vec3d operator-(const Vec3d& o) const { return {x - o.x, y - o.y, z - o.z}; }
    bool operator==(const Vec3d& o) const { return x == o.x && y == o.y && z == o.z; }
};
inline Vec3d operator*(double a, const Vec3d& v) { return {a * v.x, a * v.y, a * v.z}; }

struct Vec2d {
    double x = 0, y = 0;
};

// Row-major 3x3 (Matrix3, matrices.hpp): at(i, j) = m[3*i + j], columns are
// the basis / cell vectors.
struct Mat3d {
    double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double& at(int i, int j) { return m[3 * i + j]; }
    double at(int i, int j) const { return m[3 * i + j]; }
    Vec3d column(int j) const { return {m[j], m[3 + j], m[6 + j]}; }
    double determinant() const;
    Mat3d inverse() const;
    bool isDiagonal() const;
};

enum MinimizationCriteria { MIN_E, MIN_F, MIN_EF, NOT_SET };
enum SurfNormal { NORMAL_X, NORMAL_Y, NORMAL_Z };
enum Units { U_KCAL, U_KJ, U_EV };
enum MinimiserType { STEEPEST_DESCENT, FIRE };
enum IntegratorType { EULER, MIDPOINT, RK4 };

// InputOptions (simulation.hpp:30-60), defaults of parse.cpp:96-134.
struct InputOptions {
    std::string inputfolder, outputfolder, inputfile, xyzfile, paramfile, e_potential_file;
    std::string tipatom, dummyatom;
    Vec2d area{10, 10}, center{-1, -1};
    double dx = 0.1, dy = 0.1, dz = 0.1;
    double zlow = 6.0, zhigh = 10.0;
    bool vdw_pbc = false;
    Vec3d cell_a, cell_b, cell_c;
    SurfNormal normal = NORMAL_Z;
    Units units = U_KCAL;
    bool units_given = false;
    bool coulomb = false, tip_dummy_coulomb = false, use_external_potential = false;
    int maxsteps = 5000;
    MinimizationCriteria minterm = NOT_SET;
    double etol = 0.01, ftol = 0.01, dt = 0.001;
    int bufsize = 1000;
    bool gzip = true, statistics = false, flexible = false, rigidgrid = false;
    bool xyz_charges = false;
    MinimiserType minimiser_type = FIRE;
    IntegratorType integrator_type = MIDPOINT;
    // option values as written in the input file, for the settings echo
    std::string units_text = "kcal/mol", minterm_text, normal_text = "z";
};

struct AtomParameters {
    double eps = 0, sig = 0, q = 0, mass = 0;
};

struct OverwriteParameters {
    std::string a, b;   // unordered pair
    bool morse = false;
    double eps = 0, sig = 0, de = 0, alpha = 0, re = 0;
};

// A possible bond between two atom types (PossibleBond, interactions.hpp:42-45).
struct PossibleBond {
    std::string a, b;   // unordered pair
    double r0 = 0;      // expected bond length
};

// InteractionParameters (interactions.hpp:47-55).
struct InteractionParameters {
    double qbase = 0;
    bool qbase_set = false;
    double tip_dummy_k = 0, tip_dummy_r0 = 0;
    bool harm_set = false;
    std::map<std::string, AtomParameters> atom_parameters;
    std::vector<OverwriteParameters> overwrite_parameters;
    // flexible on (readFlexibleParameters, parse.cpp:700-782)
    double bond_k = 0, angle_k = 0, dihedral_k = 0;
    double substrate_eps = 0, substrate_sig = 0, substrate_lambda = 0, substrate_rc = 0, substrate_k = 0;
    std::vector<PossibleBond> possible_bonds;
};

// Mixing rules (interactions.hpp:58-64).
inline double mixsig(double s1, double s2) { return (s1 + s2) / 2; }
inline double mixeps(double e1, double e2) { return std::sqrt(e1 * e2); }

// Device-table builder that Interaction::lower() writes into.
struct Lowering {
    std::vector<mafm_site> sites;
    std::vector<int> site_atom;        // surface atom index of each site
    std::vector<Vec3d> site_shift;     // periodic image shift of each site
    std::vector<double> image_shifts;  // factored shifts (n_images x 3) after factorImages()
    int n_images = 0;
    mafm_tip_dummy tip_dummy{};
    bool tip_dummy_vdw_set = false;
    int force_grid = -1;
    int potential_grid = -1;
    const std::vector<Vec3d>* positions = nullptr;
};

// Rewrite a site list of the form "every atom for shift 1, every atom for
// shift 2, ..." (vdw_pbc, simulation.cpp:298-312) as one block of atoms plus
// the list of shifts, so the device streams each atom once per image.
void factorImages(Lowering& low);

// Interaction (interactions.hpp:68-76): immutable after build; lowers itself
// into the device tables.  atom indices: 0 dummy, 1 tip, >= 2 surface.
class Interaction {
  public:
    virtual ~Interaction() = default;
    virtual bool isTipSurface() const = 0;
    virtual void lower(Lowering& out) const = 0;
    // flexible on: the term of the whole-list evaluation (mafm_flex_term)
    virtual void lowerFlex(std::vector<mafm_flex_term>& out) const = 0;
};

class LJInteraction : public Interaction {
  public:
    LJInteraction(int i1, int i2, double es6, double es12, Vec3d shift = Vec3d())
        : i1_(i1), i2_(i2), es6_(es6), es12_(es12), shift_(shift) {}
    bool isTipSurface() const override { return i1_ == 1; }
    void lower(Lowering& out) const override;
    void lowerFlex(std::vector<mafm_flex_term>& out) const override;

  private:
    int i1_, i2_;
    double es6_, es12_;
    Vec3d shift_;
};

class MorseInteraction : public Interaction {
  public:
    MorseInteraction(int i1, int i2, double de, double a, double re, Vec3d shift = Vec3d())
        : i1_(i1), i2_(i2), de_(de), a_(a), re_(re), shift_(shift) {}
    bool isTipSurface() const override { return i1_ == 1; }
    void lower(Lowering& out) const override;
    void lowerFlex(std::vector<mafm_flex_term>& out) const override;

  private:
    int i1_, i2_;
    double de_, a_, re_;
    Vec3d shift_;
};

class CoulombInteraction : public Interaction {
  public:
    CoulombInteraction(int i1, int i2, double qq) : i1_(i1), i2_(i2), qq_(qq) {}
    bool isTipSurface() const override { return i1_ == 1; }
    void lower(Lowering& out) const override;
    void lowerFlex(std::vector<mafm_flex_term>& out) const override;

  private:
    int i1_, i2_;
    double qq_;
};

class TipHarmonicInteraction : public Interaction {
  public:
    TipHarmonicInteraction(int i1, int i2, double k, double r0) : i1_(i1), i2_(i2), k_(k), r0_(r0) {}
    bool isTipSurface() const override { return false; }
    void lower(Lowering& out) const override;
    void lowerFlex(std::vector<mafm_flex_term>& out) const override;

  private:
    int i1_, i2_;
    double k_, r0_;
};

// The flexible-surface terms (interactions.hpp:186-303).  They exist only
// with flexible on and lower to mafm_flex_term.
class FlexOnlyInteraction : public Interaction {
  public:
    bool isTipSurface() const override { return false; }
    void lower(Lowering& out) const override;
};

class XYHarmonicInteraction : public FlexOnlyInteraction {
  public:
    XYHarmonicInteraction(int i, double k, Vec2d p0) : i_(i), k_(k), p0_(p0) {}
    void lowerFlex(std::vector<mafm_flex_term>& out) const override;
  private:
    int i_;
    double k_;
    Vec2d p0_;
};

class SubstrateInteraction : public FlexOnlyInteraction {
  public:
    // the derived constants of the reference's constructor (interactions.hpp:222-229)
    SubstrateInteraction(int i, double eps, double sig, double rc, double lambda);
    void lowerFlex(std::vector<mafm_flex_term>& out) const override;
  private:
    int i_;
    double sig_, rc_, multiplier_, ulj_, ushift_, z0_;
};

class HarmonicInteraction : public FlexOnlyInteraction {
  public:
    HarmonicInteraction(int i1, int i2, double k, double r0) : i1_(i1), i2_(i2), k_(k), r0_(r0) {}
    void lowerFlex(std::vector<mafm_flex_term>& out) const override;
  private:
    int i1_, i2_;
    double k_, r0_;
};

class HarmonicAngleInteraction : public FlexOnlyInteraction {
  public:
    HarmonicAngleInteraction(int shared, int i1, int i2, double k, double theta0)
        : shared_(shared), i1_(i1), i2_(i2), k_(k), theta0_(theta0) {}
    void lowerFlex(std::vector<mafm_flex_term>& out) const override;
  private:
    int shared_, i1_, i2_;
    double k_, theta0_;
};

class HarmonicDihedralInteraction : public FlexOnlyInteraction {
  public:
    HarmonicDihedralInteraction(int i1, int i2, int i3, int i4, double k, double sigma0)
        : i1_(i1), i2_(i2), i3_(i3), i4_(i4), k_(k), sigma0_(sigma0) {}
    void lowerFlex(std::vector<mafm_flex_term>& out) const override;
  private:
    int i1_, i2_, i3_, i4_;
    double k_, sigma0_;
};

// ForceGrid bound to a device grid handle (force_grid.hpp:15-51).
struct ForceGrid {
    int handle = -1;
    int n[3] = {0, 0, 0};
    bool periodic = false;
    Mat3d basis;
    Vec3d offset;
};

class GridInteraction : public Interaction {
  public:
    explicit GridInteraction(const ForceGrid& g) : grid_(g) {}
    bool isTipSurface() const override { return true; }
    void lower(Lowering& out) const override { out.force_grid = grid_.handle; }
    void lowerFlex(std::vector<mafm_flex_term>& out) const override;
    const ForceGrid& grid() const { return grid_; }

  private:
    ForceGrid grid_;
};

class ElectrostaticPotentialInteraction : public Interaction {
  public:
    explicit ElectrostaticPotentialInteraction(const ForceGrid& g) : grid_(g) {}
    bool isTipSurface() const override { return true; }
    void lower(Lowering& out) const override { out.potential_grid = grid_.handle; }
    void lowerFlex(std::vector<mafm_flex_term>& out) const override;
    const ForceGrid& grid() const { return grid_; }

  private:
    ForceGrid grid_;
};

// System (system.hpp:13-77): the frozen surface plus dummy (0) and tip (1).
class System {
  public:
    void initialize(int n_atoms);
    void centerMolecule(Vec2d pos);
    void setMoleculeZ();
    void setUnitCell(const std::vector<Vec3d>& cell_vectors);
    void rotateCoordAxes(const std::string& seq);
    const Mat3d& getUnitCell() const { return cell_; }
    const Vec3d& getOffset() const { return offset_; }
    void setTipDummyDistance(double d) { tip_dummy_d_ = d; }
    double getTipDummyDistance() const { return tip_dummy_d_; }
    bool tipPbc() const { return tip_pbc_; }

    int n_atoms_ = 0;   // including tip and dummy
    std::vector<Vec3d> positions_;
    std::vector<double> charges_;
    std::vector<int> fixed_;
    std::vector<std::string> types_;

  private:
    bool tip_pbc_ = false;
    double tip_dummy_d_ = 0;
    Vec3d offset_;
    Mat3d cell_;
};

// Gaussian cube reader (cube_io.hpp:27-59).
struct Cube {
    int n[3] = {0, 0, 0};
    Vec3d voxel[3];
    Vec3d origin;
    int n_atoms = 0;
    std::vector<double> values;   // x-major, z fastest
};
Cube readCube(const std::string& path, bool with_values);
void rotateCubeZXY(Cube& c);

class Simulation {
  public:
    explicit Simulation(mafm_ctx* ctx) : ctx_(ctx) {}
    ~Simulation();
    void parseCommandLine(const std::string& input, const std::string& output);
    void readInputFile();
    void readXYZFile();
    void readParameterFile();
    void readFlexibleParameters(const std::string& path);
    void initialize();
    void buildInteractions();
    // lower the interaction list into a device system description
    void lowerSystem();
    // flexible on: lower the whole list into a term table
    void lowerFlexSystem();
    const mafm_system_desc& systemDesc() const { return desc_; }
    const mafm_flex_desc& flexDesc() const { return flex_desc_; }
    // System::makeXYZFile (system.cpp:54-65) of the state column at every
    // height of the last flexible run; the console lines go to log_
    void writeStates();
    mafm_scan_desc scanDesc(const mafm_run_opts& o) const;
    void run(const mafm_run_opts& opts, mafm_scan_stats* stats);
    void fetchRecords();
    void writeOutput(int threads);
    // the file of height k from this job's records `recs` (records_ or
    // pinned host memory); returns an error text, empty on success
    std::string writeHeight(int k, const mafm_record* recs) const;
    // the lines of height k for pixel columns [c0, c1), appended to out
    void formatLines(int k, const mafm_record* recs, size_t c0, size_t c1, std::string& out) const;
    // scan-%06.3f.dat[.gz] of height k in the output folder
    std::string heightPath(int k) const;
    // statistics.txt (mechafm.cpp:129-152) when the options ask for it
    void writeStatistics(const mafm_record* recs) const;
    // Delta f images (simple-vafm-image.py:76-204) from this job's records,
    // one deltaf-%06.3f.dat per image named by its lowest height
    void writeDeltaF(const mafm_deltaf_desc& d);
    // warnings raised while reading the inputs, then the input settings, in
    // the reference's console format (messages.cpp:29-56, parse.cpp:407-473)
    std::string settingsText() const;
    std::vector<std::string> warnings_;

    InputOptions options_;
    InteractionParameters interaction_parameters_;
    System system;
    std::vector<std::unique_ptr<Interaction>> interactions_;
    int n_points_[3] = {0, 0, 0};
    std::vector<mafm_record> records_;
    mafm_scan_stats last_stats_{};
    // when the job was opened: the wall times of statistics.txt run from here
    std::chrono::steady_clock::time_point opened_ = std::chrono::steady_clock::now();
    double grid_build_ms_ = 0;
    int force_grid_n_[3] = {0, 0, 0};
    mafm_ctx* ctx_;
    bool initialized_ = false;
    // set by batch runs: this simulation's records start here in the context,
    // and it is system `system_index_` of the launch
    size_t record_base_ = 0;
    int system_index_ = 0;
    // the context's scan count right after this simulation's scan (0: none):
    // another scan on the context replaces the records
    uint64_t scan_generation_ = 0;
    // set by mafm_batch_run_output: records_ holds this job's records, the
    // context no longer does
    bool host_records_ = false;
    void requireOwnRecords() const;
    // console lines of the output writing ("+- Wrote ..."), in order
    std::string log_;
    int n_tip_surface_terms_ = 0;

  private:
    void calculateTipDummyDistance();
    bool findOverwriteParameters(int i1, int i2, OverwriteParameters& op) const;
    void addVDWInteraction(int i1, int i2, Vec3d shift = Vec3d());
    void addCoulombInteraction(int i1, int i2);
    void buildTipSurfaceInteractions();
    void buildTipGridInteractions();
    void buildTipDummyInteractions();
    void buildSurfaceSurfaceInteractions();
    void buildSubstrateInteractions();
    ForceGrid buildElectrostaticPotential();
    std::vector<int> owned_grids_;
    Lowering low_;
    mafm_system_desc desc_{};
    std::vector<mafm_flex_term> flex_terms_;
    std::vector<double> flex_pos_;
    std::vector<int32_t> flex_fixed_;
    mafm_flex_desc flex_desc_{};
};

}  // namespace mafm
