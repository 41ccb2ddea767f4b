// ramses_reader.cpp -- RAMSES particle files (part_NNNNN.outCCCCC) read natively, one OpenMP
// thread per file: _read_ramses_new_1010 of src/input_output.py:624-760 (Ra3 and Ra4 layouts,
// families all / dm / star).  Each file is read whole and its Fortran sequential records (a
// 4-byte length before and after each payload) are parsed in place; particles are scattered
// to their dense-id slots as the reference does.  For Ra4, the stellar records after the tag
// follow the output's field descriptor when it has one (_ra4_stellar_descriptor and
// _read_ra4_stellar_record_block, :547-620): birth times, metallicities, initial masses and
// chem_<element> abundances are read where the descriptor places them, other fields skipped.
#include <omp.h>

#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ramses_reader.h"

namespace ahop {

const char *const RAMSES_CHEM[RAMSES_NCHEM] = {"H", "O", "Fe", "Mg", "C", "N", "Si", "S", "D"};

namespace {

// roles of the fields of the Ra4 stellar block
enum { F_SKIP = -1, F_BIRTH = 0, F_METAL = 1, F_M0 = 2, F_CHEM0 = 3 };

// The stellar block of a Ra4 file as the descriptor lays it out: positions start..stop of
// the descriptor's field list are the records after the tag record (start = the field after
// "tag", or the first wanted field without one; stop = the last wanted field's first
// occurrence), each read into its role or skipped.
struct Ra4Layout {
    RamsesDescriptorInfo info;
    std::vector<int> role;     // roles of positions start..stop (empty: nothing to read)
};

std::string strip(const std::string &x)
{
    size_t a = 0, b = x.size();
    while (a < b && std::isspace((unsigned char)x[a])) ++a;
    while (b > a && std::isspace((unsigned char)x[b - 1])) --b;
    return x.substr(a, b - a);
}

// the second comma-separated column of every line that is not a comment (:530-544)
bool descriptor_fields(const std::string &repository, std::vector<std::string> &fields)
{
    const std::string path = repository + "/part_file_descriptor.txt";
    std::ifstream in(path);
    if (!in) return false;
    std::string line;
    while (std::getline(in, line)) {
        if (strip(line).rfind("#", 0) == 0) continue;
        std::vector<std::string> cols;
        size_t p0 = 0;
        for (;;) {
            const size_t c = line.find(',', p0);
            cols.push_back(strip(line.substr(p0, c == std::string::npos ? std::string::npos : c - p0)));
            if (c == std::string::npos) break;
            p0 = c + 1;
        }
        if (cols.size() >= 2) fields.push_back(cols[1]);
    }
    return true;
}

int field_role(const std::string &name)
{
    if (name == "birth_time") return F_BIRTH;
    if (name == "metallicity") return F_METAL;
    if (name == "initial_mass") return F_M0;
    for (int e = 0; e < RAMSES_NCHEM; ++e)
        if (name == std::string("chem_") + RAMSES_CHEM[e]) return F_CHEM0 + e;
    return F_SKIP;
}

Ra4Layout ra4_layout(const std::string &repository)
{
    Ra4Layout L;
    std::vector<std::string> fields;
    if (!descriptor_fields(repository, fields)) return L;
    L.info.present = true;
    const std::string path = repository + "/part_file_descriptor.txt";
    auto first = [&](const std::string &name) {
        for (size_t i = 0; i < fields.size(); ++i)
            if (fields[i] == name) return (int)i;
        return -1;
    };
    const int imetal = first("metallicity"), im0 = first("initial_mass");
    if (im0 >= 0 && imetal < 0)
        throw std::runtime_error("Ra4 descriptor has initial_mass but no metallicity: " + path);
    if (im0 >= 0 && im0 <= imetal)
        throw std::runtime_error("Unsupported Ra4 initial_mass record order: " + path);
    L.info.has_m0 = im0 >= 0;
    int lo = -1, hi = -1;
    for (size_t i = 0; i < fields.size(); ++i) {
        const int r = field_role(fields[i]);
        if (r == F_SKIP || first(fields[i]) != (int)i) continue;
        if (r >= F_CHEM0) L.info.chem_mask |= 1 << (r - F_CHEM0);
        lo = lo < 0 ? (int)i : std::min(lo, (int)i);
        hi = std::max(hi, (int)i);
    }
    if (hi < 0) return L;                       // no stellar field listed: nothing is read
    const int itag = first("tag");
    const int start = itag >= 0 ? itag + 1 : lo;
    for (int i = start; i <= hi; ++i) L.role.push_back(field_role(fields[i]));
    return L;
}

struct Records {
    std::vector<char> buf;
    size_t pos = 0;
    std::string name;

    void load(const std::string &path)
    {
        name = path;
        FILE *f = std::fopen(path.c_str(), "rb");
        if (!f) throw std::runtime_error("cannot open " + path);
        std::fseek(f, 0, SEEK_END);
        const long n = std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        buf.resize(n > 0 ? (size_t)n : 0);
        const size_t got = n > 0 ? std::fread(buf.data(), 1, (size_t)n, f) : 0;
        std::fclose(f);
        if (got != buf.size()) throw std::runtime_error("short read of " + path);
        pos = 0;
    }
    // the next record's payload: pointer and byte length (markers checked)
    const char *next(size_t &len)
    {
        if (pos + 4 > buf.size()) throw std::runtime_error(name + ": record past the end");
        uint32_t a, b;
        std::memcpy(&a, buf.data() + pos, 4);
        if (pos + 8 + (size_t)a > buf.size()) throw std::runtime_error(name + ": record past the end");
        std::memcpy(&b, buf.data() + pos + 4 + a, 4);
        if (a != b) throw std::runtime_error(name + ": record markers disagree");
        const char *p = buf.data() + pos + 4;
        pos += 8 + (size_t)a;
        len = a;
        return p;
    }
    void skip(int n = 1)
    {
        size_t len;
        for (int i = 0; i < n; ++i) (void)next(len);
    }
    int32_t one_int()
    {
        size_t len;
        const char *p = next(len);
        if (len < 4) throw std::runtime_error(name + ": short integer record");
        int32_t v;
        std::memcpy(&v, p, 4);
        return v;
    }
    template <typename T>
    const T *array(int64_t n)
    {
        size_t len;
        const char *p = next(len);
        if (len != sizeof(T) * (size_t)n) throw std::runtime_error(name + ": record of unexpected size");
        return reinterpret_cast<const T *>(p);   // payloads are 4-byte aligned at most: read by memcpy
    }
};

template <typename T>
inline T at(const T *p, int64_t i)
{
    T v;
    std::memcpy(&v, reinterpret_cast<const char *>(p) + sizeof(T) * i, sizeof(T));
    return v;
}

std::string part_name(const std::string &repo, const std::string &nchar, int icpu)
{
    char tail[32];
    std::snprintf(tail, sizeof tail, ".out%05d", icpu);
    return repo + "/part_" + nchar + tail;
}

// one file: classify its particles; with outputs, scatter the selected ones
void one_file(const RamsesArgs &A, const Ra4Layout &L, int icpu, RamsesFileCount &cnt,
              bool count_only)
{
    Records f;
    f.load(part_name(A.repository, A.nchar, icpu));
    (void)f.one_int();                  // ncpu
    const int ndim = f.one_int();
    const int64_t npart = f.one_int();
    f.skip(1);                          // localseed
    const int nstar = f.one_int();
    f.skip(2);                          // mstar_tot, mstar_lost
    const int nsink = f.one_int();
    if (ndim != 3) throw std::runtime_error(f.name + ": ndim must be 3");
    const double *xp[3] = {nullptr, nullptr, nullptr}, *vp[3] = {nullptr, nullptr, nullptr};
    const double *mp = nullptr;
    if (count_only) {
        f.skip(2 * ndim + 1);
    } else {
        for (int d = 0; d < 3; ++d) xp[d] = f.array<double>(npart);
        for (int d = 0; d < 3; ++d) vp[d] = f.array<double>(npart);
        mp = f.array<double>(npart);
    }
    const int32_t *idp = f.array<int32_t>(npart);
    f.skip(1);                          // level
    const int8_t *fam = nullptr;
    if (A.rver == 4) {
        fam = f.array<int8_t>(npart);
        f.skip(1);                      // tag
    }
    const double *birth = nullptr, *metal = nullptr, *m0 = nullptr;
    const double *chem[RAMSES_NCHEM] = {};
    if (nstar > 0 || nsink > 0) {
        const bool keep = A.stellar && !count_only && A.family != RAMSES_DM;
        if (A.rver == 4) {
            if (keep && !L.info.present) {
                birth = f.array<double>(npart);
                metal = f.array<double>(npart);
            } else if (keep) {
                for (const int r : L.role) {
                    if (r == F_SKIP) {
                        f.skip(1);
                        continue;
                    }
                    const double *v = f.array<double>(npart);
                    if (r == F_BIRTH) birth = v;
                    else if (r == F_METAL) metal = v;
                    else if (r == F_M0) m0 = v;
                    else chem[r - F_CHEM0] = v;
                }
            }
        } else {
            birth = f.array<double>(npart);   // Ra3 classifies by birth time
            if (keep) metal = f.array<double>(npart);
        }
    }
    auto is_dm = [&](int64_t i) {
        if (A.rver == 4) return at(fam, i) == 1;
        const double t = birth ? at(birth, i) : 0.0;
        return at(idp, i) > 0 && t == 0.0;
    };
    auto is_star = [&](int64_t i) {
        if (A.rver == 4) return at(fam, i) == 2;
        const double t = birth ? at(birth, i) : 0.0;
        const int32_t id = at(idp, i);
        return (t < 0 && id > 0) || (t != 0 && id < 0);
    };
    int64_t ndm = 0, ns = 0, nsel = 0;
    for (int64_t i = 0; i < npart; ++i) {
        const bool d = is_dm(i), s = is_star(i);
        ndm += d;
        ns += s;
        nsel += A.family == RAMSES_ALL ? (d || s) : A.family == RAMSES_DM ? d : s;
    }
    cnt.ndm = ndm;
    cnt.nstar = ns;
    cnt.nselected = nsel;
    if (count_only) return;
    // dense ids (:715-730): stars after the dark matter with family 'all'
    std::vector<int64_t> seen;
    seen.reserve(nsel);
    for (int64_t i = 0; i < npart; ++i) {
        const bool d = is_dm(i), s = is_star(i);
        const bool sel = A.family == RAMSES_ALL ? (d || s) : A.family == RAMSES_DM ? d : s;
        if (!sel) continue;
        const int64_t src = std::llabs((long long)at(idp, i));
        const int64_t local = (A.family == RAMSES_ALL && s) ? src + A.ndm_total : src;
        const int64_t ind = local - 1;
        if (ind < 0 || ind >= A.nselected_total)
            throw std::runtime_error(f.name + ": particle id outside the dense range");
        seen.push_back(ind);
        for (int d3 = 0; d3 < 3; ++d3) {
            A.pos[3 * ind + d3] = at(xp[d3], i) - 0.5;
            A.vel[3 * ind + d3] = at(vp[d3], i) * A.scale_l / A.scale_t * 1e-5;
        }
        A.mass[ind] = at(mp, i);
        if (A.family != RAMSES_ALL && A.source_pid) A.source_pid[ind] = src;
        if (s && A.stellar && A.family != RAMSES_DM && A.birth) {
            if (!birth || !metal)
                throw std::runtime_error(f.name + ": stars without a birth_time or metallicity "
                                                  "record in the field descriptor");
            const int64_t slot = src - 1;   // star_slots (:746)
            if (slot < 0 || slot >= A.nstar_total)
                throw std::runtime_error(f.name + ": star id outside the star range");
            A.birth[slot] = at(birth, i);
            if (A.metal) A.metal[slot] = at(metal, i);
            if (m0 && A.m0) A.m0[slot] = at(m0, i);
            if (A.chem)
                for (int e = 0; e < RAMSES_NCHEM; ++e)
                    if (chem[e]) A.chem[slot * RAMSES_NCHEM + e] = at(chem[e], i);
        }
    }
    // duplicates within the file are an error (:720-723)
    std::vector<int64_t> s2(seen);
    std::sort(s2.begin(), s2.end());
    for (size_t k = 1; k < s2.size(); ++k)
        if (s2[k] == s2[k - 1]) throw std::runtime_error(f.name + ": duplicate particle id");
}

void run_files(const RamsesArgs &A, RamsesFileCount *per_file, bool count_only)
{
    // every file records its own error; the lowest-numbered file's is reported, as the
    // reference, reading the files in order, would stop there
    std::vector<std::string> err(A.ncpu);
    // the descriptor is read (and checked) before any file, as read_ramses_new_101 does (:776)
    const Ra4Layout L = A.rver == 4 ? ra4_layout(A.repository) : Ra4Layout();
    const int nt = A.nthreads > 0 ? A.nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int icpu = 1; icpu <= A.ncpu; ++icpu) {
        try {
            one_file(A, L, icpu, per_file[icpu - 1], count_only);
        } catch (const std::exception &ex) {
            err[icpu - 1] = ex.what();
        }
    }
    for (const std::string &e : err)
        if (!e.empty()) throw std::runtime_error(e);
}

}  // namespace

RamsesDescriptorInfo ramses_descriptor(const std::string &repository)
{
    return ra4_layout(repository).info;
}

void ramses_count(const RamsesArgs &A, RamsesFileCount *per_file)
{
    if (A.ncpu < 1) throw std::runtime_error("ncpu must be >= 1");
    if (A.rver != 3 && A.rver != 4) throw std::runtime_error("rver must be 3 (Ra3) or 4 (Ra4)");
    run_files(A, per_file, true);
}

void ramses_read(const RamsesArgs &A, RamsesFileCount *per_file)
{
    if (A.ncpu < 1) throw std::runtime_error("ncpu must be >= 1");
    if (A.rver != 3 && A.rver != 4) throw std::runtime_error("rver must be 3 (Ra3) or 4 (Ra4)");
    if (!A.pos || !A.vel || !A.mass) throw std::runtime_error("null output array");
    run_files(A, per_file, false);
}

}  // namespace ahop
