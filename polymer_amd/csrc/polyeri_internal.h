// polyeri_internal.h -- host-side data structures of libpolyeri.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "eri_engine.h"

namespace polyeri {

// The reference's basis arrays (module BASISSET, src/pmodules.f90:64-81),
// 0-based, CC column-major ncgs x npgs.
struct RefBasis {
    int natom = 0;
    std::vector<int> iatom;
    std::vector<double> ax, ay, az;
    int npgs = 0, ncgs = 0, npshell = 0;
    std::vector<int> pax, pay, paz, pgs_atom;
    std::vector<double> pgsx, pgsy, pgsz, zt, cc;
    std::vector<double> cgsx, cgsy, cgsz;
    std::vector<int> p_shell;       // npshell x 125, -1 = empty
    std::vector<int> p_shell_ang;
};

// A contracted Cartesian shell of this library.  Component mu of the shell
// is chi = sum_p coef[p] (x-A)^lx (y-A)^ly (z-A)^lz exp(-alpha[p] r^2) and
// the reference's contracted function phi_c = scale * chi.
struct Shell {
    int L = 0;
    int atom = 0;
    double x = 0, y = 0, z = 0;
    std::vector<double> alpha, coef;
    int f0 = 0;                     // first internal basis-function index
};

struct PairInfo {
    int a, b;                       // shells, L_a >= L_b
    int cls;                        // pair class (La,Lb) index
    int prim_off, nprim;            // into the PrimPair array
    double ABx, ABy, ABz;
};

// Basis as this library lays it out: shells sorted by L (d, p, s), every
// contracted function mapped to the reference's index and scale.
struct IntBasis {
    std::vector<Shell> shells;
    int nbf = 0;
    std::vector<int> ref_of_int;    // nbf -> ncgs index
    std::vector<double> scale;      // nbf: phi_ref = scale * chi
    std::vector<int> int_of_ref;    // ncgs -> nbf or -1 (ghost)
    std::vector<int> natom_z;       // atomic numbers
    std::vector<double> atom_xyz;   // 3 per atom
};

inline int pair_class(int la, int lb) { return la * (la + 1) / 2 + lb; }
constexpr int NPAIRCLASS = 6;       // (ss) (ps) (pp) (ds) (dp) (dd)
constexpr int MAXL = 2;

// Screened work list of one class pair (X >= Y): bra rows are the pairs of
// class X in Schwarz order; the kets of class Y are grouped in buckets of
// equal primitive-pair count (Schwarz order inside each bucket); row i holds,
// bucket by bucket, the kets j with Q_i Q_j >= tol (and rank_j <= rank_i
// when X == Y), each a prefix of its bucket.  With a mixed-precision cut
// the class pair has two lists: prec 0 holds the kets with Q_i Q_j >= cut
// (a prefix of each bucket), prec 1 those with tol <= Q_i Q_j < cut (a
// range of each bucket that starts at rf[row][bucket]).
struct WorkList {
    int X = 0, Y = 0, nrows = 0, nbucket = 1;
    int prec = 0;                   // 0: FP64, 1: FP32
    long long nwork = 0;
    std::vector<long long> row;     // nrows + 1, exclusive prefix
    std::vector<int> rb;            // nrows x (nbucket + 1), offsets inside a row
    std::vector<int> rf;            // prec 1: nrows x nbucket, first ket inside each bucket
    double prim_quartets = 0;
};

// Host functions (host.cpp).
int read_basis(const std::string& path, int natom, const int* iatom, const double* xyz,
               RefBasis& out, std::string& err);
int extract_shells(const RefBasis& rb, IntBasis& ib, std::string& err);
void build_pairs(const IntBasis& ib, double primtol, std::vector<PairInfo>& pairs,
                 std::vector<PrimPair>& prims);
void boys_table(std::vector<double>& tab);       // BOYS_NT x BOYS_NM
// Generally contracted pairs for the classes (ss), (ps), (ds).  Shell
// groups: an s shell whose exponents all occur among those of an earlier s
// shell on the same centre joins that shell's group (two shells at most);
// every other shell is a group of its own.  A pair of groups g1 <= g2 (the
// second of L 0) has as members the shell pairs of the two groups
// (unordered when g1 == g2), primitive pairs over the exponents of the
// groups' first shells (K without coefficients) and one weight per member
// and primitive pair.  member_pair[4 i + j]: index in pairs of member j of
// gpairs[i] (-1: no such pair, or j >= m).  Returns the number of two-shell
// groups; with none the outputs stay empty.
int build_gc_pairs(const IntBasis& ib, double primtol, const std::vector<PairInfo>& pairs,
                   std::vector<PairInfo>& gpairs, std::vector<GcMeta>& gmeta, std::vector<GcPrim>& gprims,
                   std::vector<int>& member_pair);
// class_ids[X] must hold the pairs of class X in Schwarz order (bra rows);
// fills ket_ids / bstart (bucketed ket order) and the lists.  cut > tol
// splits a class pair into an FP64 list and an FP32 list (see WorkList) when
// bit X * NPAIRCLASS + Y of split_mask is set (every class pair by default).
// nosym: every ket of class Y for every bra of class X (no rank_j <= rank_i
// restriction), for the periodic lattice sums.
void build_work_lists(const std::vector<PairInfo>& pairs, const std::vector<double>& Q, double tol,
                      double cut, const std::vector<int> class_ids[], std::vector<int> ket_ids[],
                      std::vector<int> bstart[], std::vector<WorkList>& out, bool nosym = false,
                      uint64_t split_mask = ~0ull);
struct int2_host {
    int x, y;
};
// Periodic pairs (cell offset Q in [-rc1, rc1] of the second shell), cells[i]
// = {Q, swap}: swap = 1 when the stored first shell is the one in cell Q.
void build_pairs_cells(const IntBasis& ib, double primtol, double period, int rc1, std::vector<PairInfo>& pairs,
                       std::vector<PrimPair>& prims, std::vector<int2_host>& cells);
// S, T, V in the internal ordering (nbf x nbf, row-major, symmetric).
void one_electron(const IntBasis& ib, const std::vector<double>& boys, std::vector<double>& S,
                  std::vector<double>& T, std::vector<double>& V);
double nuclear_repulsion(const IntBasis& ib);
// One-electron and nuclear-repulsion parts of the nuclear gradient at fixed
// D and W (library basis, nbf x nbf), grad (3 natom) overwritten.
void one_electron_gradient(const IntBasis& ib, const std::vector<int>& shell_atom, const std::vector<double>& boys,
                           const double* D, const double* W, double* grad);
// Cartesian exponents of component k of a shell of angular momentum l.
void cart_comp(int l, int k, int& lx, int& ly, int& lz);

}  // namespace polyeri
