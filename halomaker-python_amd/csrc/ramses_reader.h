// ramses_reader.h -- RAMSES particle files read natively (ramses_reader.cpp).
#pragma once
#include <cstdint>
#include <string>

namespace ahop {

enum RamsesFamily { RAMSES_ALL = 0, RAMSES_DM = 1, RAMSES_STAR = 2 };

struct RamsesFileCount {
    int64_t ndm = 0, nstar = 0, nselected = 0;
};

struct RamsesArgs {
    std::string repository, nchar;   // output directory and its 5-character number
    int ncpu = 0, rver = 3, family = RAMSES_ALL, nthreads = 0;
    bool stellar = true;             // stellar payload: birth epochs and metallicities
    // read pass only
    int64_t ndm_total = 0, nselected_total = 0, nstar_total = 0;
    double scale_l = 1.0, scale_t = 1.0;
    double *pos = nullptr, *vel = nullptr, *mass = nullptr;   // nselected x 3, x 3, x 1
    int64_t *source_pid = nullptr;                            // family dm / star
    double *birth = nullptr, *metal = nullptr;                // nstar slots
    // Ra4 with a field descriptor: initial masses (nstar slots) and abundances (nstar x
    // RAMSES_NCHEM, in RAMSES_CHEM order); entries the files do not provide are left as they are
    double *m0 = nullptr, *chem = nullptr;
};

// the chemical elements a Ra4 descriptor may list as chem_<element> (src/chem_species.py:1)
constexpr int RAMSES_NCHEM = 9;
extern const char *const RAMSES_CHEM[RAMSES_NCHEM];

// What the Ra4 field descriptor (part_file_descriptor.txt) of an output provides for stars:
// present (a descriptor was found), initial masses, and a bit per RAMSES_CHEM element.
struct RamsesDescriptorInfo {
    bool present = false, has_m0 = false;
    int chem_mask = 0;
};
RamsesDescriptorInfo ramses_descriptor(const std::string &repository);

void ramses_count(const RamsesArgs &A, RamsesFileCount *per_file);
void ramses_read(const RamsesArgs &A, RamsesFileCount *per_file);

}  // namespace ahop
