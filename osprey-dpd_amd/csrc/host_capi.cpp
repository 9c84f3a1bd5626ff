// host_capi.cpp -- C ABI of the host layer (include/dpd_host.h): dmpci input, starting
// states, restart snapshots and the CSimBox::Run-shaped driver over the engine.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "dmpci.hpp"
#include "dpd_engine.h"
#include "dpd_host.h"

using dpd::host::InputData;
using dpd::host::System;

struct dpd_input {
    InputData d;
    System s;
    std::vector<double> a, g;
};

namespace {

void put_err(char* err, int32_t len, const std::string& msg)
{
    if (err && len > 0) {
        std::strncpy(err, msg.c_str(), (size_t)len - 1);
        err[len - 1] = 0;
    }
}

int finish_input(dpd_input* in, dpd_input** out, char* err, int32_t errlen)
{
    std::string e;
    if (!dpd::host::build_system(in->d, in->s, e)) {
        put_err(err, errlen, e);
        delete in;
        return DPD_ERR_IO;
    }
    in->a = in->d.a;
    in->g = in->d.g;
    *out = in;
    return DPD_OK;
}

double box_len(const InputData& d, int a)
{
    return a == 0 ? d.nx * d.wx : (a == 1 ? d.ny * d.wy : d.nz * d.wz);
}

// The reference's uniform stream for building states: its 64-bit LCG (CCNTCell::Randf,
// src/CNTCell.cpp:2454-2465) started from |RNGSeed| (CCNTCell::SetRNGSeed, :134-137)
struct RefUniform {
    uint64_t state;
    double next()
    {
        state = 6364136223846793005ull * state + 1ull;
        return (double)(uint32_t)(state >> 32) * (1.0 / 4294967296.0);
    }
};

// CBuilder::MaxwellDist (src/Builder.cpp:104-129): a table of max(n, 10^7) speeds, each
// from the factor the previous one left; the last entry repeats the one before it.  The
// table depends on its size alone, so the last one built is kept (10^7 entries take ~0.5 s,
// and every state of up to 10^7 beads uses the same table).
std::shared_ptr<const std::vector<double>> maxwell_table(int64_t n)
{
    static std::mutex mu;
    static std::shared_ptr<const std::vector<double>> cache;
    const int64_t m = std::max<int64_t>(n, 10000000);
    std::lock_guard<std::mutex> lock(mu);
    if (cache && (int64_t)cache->size() == m) return cache;
    auto t = std::make_shared<std::vector<double>>((size_t)m, 0.0);
    const double dm = (double)m;
    double factor = 1.0;
    for (int64_t j = 0; j < m - 1; ++j) {
        const double vel = std::sqrt(2.0 * std::log(dm * factor / (dm - factor)));
        factor = std::exp(0.5 * vel * vel);
        (*t)[(size_t)j] = vel;
    }
    (*t)[(size_t)(m - 1)] = (*t)[(size_t)(m - 2)];
    cache = t;
    return cache;
}

// CRandomBuilder::Assemble (src/RandomBuilder.cpp:71-425) without walls: per bead, in
// creation order, a speed drawn from the Maxwell table, a position uniform in the box
// (a coordinate within 1e-4 of the upper face set to 0) and a direction uniform on the
// sphere; then, bond by bond in each polymer, a tail further than 1 from its head along an
// axis is moved to the head plus that axis' component of the unit head-tail vector; then
// the centre-of-mass velocity is removed and the velocities scaled to kT exactly.
// Wrapped, unwrapped and initial positions are the same.
void build_reference_random(const dpd_input* in, uint64_t seed, double* x, double* v)
{
    const InputData& d = in->d;
    const System& s = in->s;
    const int64_t n = (int64_t)s.id.size();
    RefUniform U{seed};
    const auto tab = maxwell_table(n);
    const std::vector<double>& table = *tab;
    const double L[3] = {box_len(d, 0), box_len(d, 1), box_len(d, 2)};
    double vcm[3] = {0.0, 0.0, 0.0};
    const double two_pi = 2.0 * 3.141592653589793238;
    for (int64_t i = 0; i < n; ++i) {
        const size_t index = (size_t)((double)table.size() * U.next());
        const double vmag = table[index];
        double xp[3];
        xp[0] = L[0] * U.next();
        xp[1] = L[1] * U.next();
        xp[2] = L[2] * U.next();
        for (int a = 0; a < 3; ++a)
            if (L[a] - xp[a] < 0.0001) xp[a] = 0.0;
        const double vtheta = std::acos(1.0 - 2.0 * U.next());
        const double vphi = two_pi * U.next();
        const double vp[3] = {vmag * std::sin(vtheta) * std::cos(vphi), vmag * std::sin(vtheta) * std::sin(vphi),
                              vmag * std::cos(vtheta)};
        for (int a = 0; a < 3; ++a) {
            vcm[a] += vp[a];
            v[3 * i + a] = vp[a];
            x[3 * i + a] = xp[a];
        }
    }
    // bonds in polymer order are the engine's bond order (polymers in creation order)
    for (size_t b = 0; b < s.bond_head.size(); ++b) {
        const int64_t h = s.bond_head[b], t = s.bond_tail[b];
        double rd[3];
        for (int a = 0; a < 3; ++a) rd[a] = x[3 * t + a] - x[3 * h + a];
        const double rmag = std::sqrt(rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2]);
        for (int a = 0; a < 3; ++a)
            if (std::fabs(rd[a]) > 1.0) x[3 * t + a] = x[3 * h + a] + rd[a] / rmag;
    }
    for (int a = 0; a < 3; ++a) vcm[a] /= (double)n;
    double vtotal = 0.0;
    for (int64_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            v[3 * i + a] -= vcm[a];
            vtotal += v[3 * i + a] * v[3 * i + a];
        }
    vtotal = std::sqrt(vtotal / (double)(3 * n));
    const double sk = std::sqrt(d.kT);
    for (int64_t i = 0; i < 3 * n; ++i) v[i] = sk * v[i] / vtotal;
}

// CLamellaBuilder::Assemble (src/LamellaBuilder.cpp:126-1003) without walls or polymerised
// lipids, with the lamella set-up of CInitialStateLamella::ValidateData
// (src/InitialStateLamella.cpp:672-769):
// * the lipids (the named polymer type, in creation order) split into an upper monolayer of
//   (long)(UpperFraction n) and a lower one of the rest, each on a hexagonal lattice of
//   spacing 0.999 sqrt(1.1547 A / n_m) started 0.001 of a lattice vector from the origin;
//   rows of the lattice advance when a site passes the in-plane length plus the row's shift,
//   and a site past the length alone wraps back by it;
// * per lipid the tail (last bead) at centre +- 0.2, the head (first bead) at
//   centre +- Thickness/2, the other beads in order on the line between at spacing
//   Thickness / (2 (size - 1)) (Linearise 1), or (Linearise 0) at lattice site + 0.5 e1 U,
//   + 0.5 e2y U and centre +- Thickness/2 U along the normal, pulled back by half a lattice
//   vector past the box;
// * every other polymer (CBuilder::isRandom, src/Builder.cpp:293-490) in one of the two free
//   slabs, chosen with probability proportional to its width, beads uniform in it along the
//   normal and in 0.95 of the box across it, then the bond fix-up of the random builder;
// * unwrapped = wrapped; 2-d Maxwell velocities in the plane (CBuilder::AssignMaxwellVelocities,
//   src/Builder.cpp:4207-4383): per bead a speed from the Maxwell table and an angle, the
//   centre-of-mass velocity removed and the in-plane kinetic energy scaled to kT per
//   component.
// The uniform stream is the reference's LCG from |RNGSeed|, as for the random builder.
bool build_reference_lamella(const dpd_input* in, uint64_t seed, double* x, double* v, std::string& err)
{
    const InputData& d = in->d;
    const System& s = in->s;
    const int64_t n = (int64_t)s.id.size();
    std::vector<std::string> pol = d.state_values("Polymer", 1);
    std::vector<std::string> nrm = d.state_values("Normal", 3);
    std::vector<std::string> cen = d.state_values("Centre", 1);
    std::vector<std::string> thk = d.state_values("Thickness", 1);
    std::vector<std::string> lin = d.state_values("Linearise", 1);
    std::vector<std::string> upf = d.state_values("UpperFraction", 1);
    std::vector<std::string> plm = d.state_values("Polymerise", 1);
    if (pol.empty() || nrm.size() != 3 || cen.empty() || thk.empty() || lin.empty() || upf.empty()) {
        err = "lamella state needs Polymer, Normal, Centre, Thickness, Linearise and UpperFraction";
        return false;
    }
    if (!plm.empty() && std::atoi(plm[0].c_str()) != 0) {
        err = "polymerised lamellae are not built";
        return false;
    }
    const int X = std::atoi(nrm[0].c_str()), Y = std::atoi(nrm[1].c_str()), Z = std::atoi(nrm[2].c_str());
    if (X + Y + Z != 1 || X < 0 || Y < 0 || Z < 0) {
        err = "lamella normal must be a unit Cartesian axis";
        return false;
    }
    int ptype = -1;
    for (size_t p = 0; p < d.polymers.size(); ++p)
        if (d.polymers[p].name == pol[0]) ptype = (int)p;
    if (ptype < 0) {
        err = "lamella polymer not found";
        return false;
    }
    const double L[3] = {box_len(d, 0), box_len(d, 1), box_len(d, 2)};
    const int na = X ? 0 : (Y ? 1 : 2);                // normal axis
    const double l1 = X ? L[1] : L[0];                 // in-plane lengths
    const double l2 = Z ? L[1] : L[2];
    const double thickness = std::atof(thk[0].c_str());
    const double centre = std::atof(cen[0].c_str()) * L[na];
    if (centre + 0.5 * thickness > L[na] || centre - 0.5 * thickness < 0.0) {
        err = "lamella outside the box";
        return false;
    }
    const double area = l1 * l2;
    double lowerHead = centre - 0.5 * thickness, upperHead = centre + 0.5 * thickness;
    const bool linearise = std::atoi(lin[0].c_str()) != 0;
    const double upperFraction = std::atof(upf[0].c_str());

    // polymers in creation order: first bead and bead count
    std::vector<int64_t> pstart, plen;
    std::vector<int> ptyp;
    for (int64_t i = 0; i < n;) {
        int64_t j = i;
        while (j < n && s.poly_id[j] == s.poly_id[i]) ++j;
        pstart.push_back(i);
        plen.push_back(j - i);
        ptyp.push_back(s.poly_type[i]);
        i = j;
    }
    std::vector<size_t> lipids;
    for (size_t p = 0; p < pstart.size(); ++p)
        if (ptyp[p] == ptype) lipids.push_back(p);
    const int64_t total = (int64_t)lipids.size();
    if (total == 0) {
        err = "no lamella polymers";
        return false;
    }
    const int64_t size = plen[lipids[0]];
    if (size < 2) {
        err = "lamella polymers need a head and a tail";
        return false;
    }

    int64_t mtotal[2], first[2] = {0, 0}, lattice[2];
    double sign[2] = {1.0, -1.0}, mhead[2];
    mtotal[0] = (int64_t)(upperFraction * (double)total);
    mtotal[1] = total - mtotal[0];
    if (mtotal[0] == 0) {
        lattice[0] = 0; lattice[1] = mtotal[1];
        mhead[0] = 0.0; mhead[1] = lowerHead;
        upperHead = centre;
    } else if (mtotal[1] == 0) {
        lattice[0] = mtotal[0]; lattice[1] = 0;
        mhead[0] = upperHead; mhead[1] = 0.0;
        lowerHead = centre;
    } else {
        first[1] = mtotal[0];
        lattice[0] = mtotal[0]; lattice[1] = mtotal[1];
        mhead[0] = upperHead; mhead[1] = lowerHead;
    }
    double hs[2], e1x[2], e2x[2], e2y[2];
    for (int m = 0; m < 2; ++m) {
        hs[m] = lattice[m] > 0 ? 0.999 * std::sqrt(1.1547 * area / (double)lattice[m]) : 0.0;
        e1x[m] = hs[m];
        e2x[m] = 0.5 * hs[m];
        e2y[m] = 0.866025 * hs[m];
    }
    // head first and tail last in the bead order: every bead counts
    const double caxis = 0.5 * thickness / (double)(size - 1);

    RefUniform U{seed};
    // in-plane (p0, p1) and normal coordinate to x, y, z
    auto put = [&](int64_t b, double p0, double p1, double pn) {
        double* q = x + 3 * b;
        if (X) { q[0] = pn; q[1] = p0; q[2] = p1; }
        else if (Y) { q[0] = p1; q[1] = pn; q[2] = p0; }
        else { q[0] = p0; q[1] = p1; q[2] = pn; }
    };
    for (int m = 0; m < 2; ++m) {
        double hexp[2] = {0.001 * e1x[m], 0.001 * e2y[m]};
        double xp[2] = {hexp[0], hexp[1]};
        int64_t rowCounter = 0;
        for (int64_t iv = first[m]; iv < first[m] + lattice[m]; ++iv) {
            const size_t p = lipids[(size_t)iv];
            const int64_t b0 = pstart[p], b1 = b0 + plen[p];
            const int64_t head = b0, tail = b1 - 1;
            put(tail, xp[0], xp[1], centre + sign[m] * 0.2);
            put(head, xp[0], xp[1], mhead[m]);
            double beadCounter = 1.0;
            for (int64_t b = b0; b < b1; ++b) {
                if (b == head || b == tail) continue;
                if (linearise) {
                    put(b, xp[0], xp[1], mhead[m] - sign[m] * beadCounter * caxis);
                    beadCounter += 1.0;
                } else {
                    // the draws in the reference's order of the three coordinates
                    double q[3];
                    if (X) {
                        q[0] = centre + 0.5 * sign[m] * thickness * U.next();
                        q[1] = xp[0] + 0.5 * e1x[m] * U.next();
                        q[2] = xp[1] + 0.5 * e2y[m] * U.next();
                        if (q[1] > L[1]) q[1] -= 0.5 * e1x[m];
                        if (q[2] > L[2]) q[2] -= 0.5 * e2y[m];
                    } else if (Y) {
                        q[0] = xp[1] + 0.5 * e2y[m] * U.next();
                        q[1] = centre + 0.5 * sign[m] * thickness * U.next();
                        q[2] = xp[0] + 0.5 * e1x[m] * U.next();
                        if (q[0] > L[0]) q[0] -= 0.5 * e2y[m];
                        if (q[2] > L[2]) q[2] -= 0.5 * e1x[m];
                    } else {
                        q[0] = xp[0] + 0.5 * e1x[m] * U.next();
                        q[1] = xp[1] + 0.5 * e2y[m] * U.next();
                        q[2] = centre + 0.5 * sign[m] * thickness * U.next();
                        if (q[0] > L[0]) q[0] -= 0.5 * e1x[m];
                        if (q[1] > L[1]) q[1] -= 0.5 * e2y[m];
                    }
                    for (int a = 0; a < 3; ++a) x[3 * b + a] = q[a];
                }
            }
            hexp[0] += e1x[m];
            if (hexp[0] > l1 + (double)rowCounter * e2x[m]) {
                ++rowCounter;
                hexp[0] = (double)rowCounter * e2x[m];
                hexp[1] += e2y[m];
                xp[0] = hexp[0];
                xp[1] = hexp[1];
                if (hexp[1] >= l2) {
                    err = "lamella lattice outside SimBox";
                    return false;
                }
            } else if (hexp[0] > l1) {
                xp[0] = hexp[0] - l1;
            } else {
                xp[0] = hexp[0];
                xp[1] = hexp[1];
            }
        }
    }

    // the other polymers in the two free slabs (no walls: [0, lowerHead], [upperHead, L])
    const double ex[6] = {0.0, 0.0, lowerHead, upperHead, L[na], L[na]};
    const double freeLength = (ex[2] - ex[1]) + (ex[4] - ex[3]);
    const double cum[2] = {(ex[2] - ex[1]) / freeLength, (ex[2] - ex[1]) / freeLength + (ex[4] - ex[3]) / freeLength};
    for (size_t p = 0; p < pstart.size(); ++p) {
        if (ptyp[p] == ptype) continue;
        const double r = U.next();
        int index = 0;
        bool unassigned = true;
        for (int k = 0; k < 2; ++k)
            if (unassigned && r <= cum[k]) { unassigned = false; index = k; }
        const double width = ex[2 * index + 2] - ex[2 * index + 1];
        const double base = ex[2 * index + 1];
        for (int64_t b = pstart[p]; b < pstart[p] + plen[p]; ++b) {
            double q[3];
            if (X) {
                q[0] = base + width * U.next();
                q[1] = 0.95 * L[1] * U.next();
                q[2] = 0.95 * L[2] * U.next();
            } else if (Y) {
                q[0] = 0.95 * L[0] * U.next();
                q[1] = base + width * U.next();
                q[2] = 0.95 * L[2] * U.next();
            } else {
                q[0] = 0.95 * L[0] * U.next();
                q[1] = 0.95 * L[1] * U.next();
                q[2] = base + width * U.next();
            }
            for (int a = 0; a < 3; ++a) x[3 * b + a] = q[a];
        }
    }
    // the bond fix-up of the free polymers (their bonds, in polymer order)
    for (size_t b = 0; b < s.bond_head.size(); ++b) {
        const int64_t h = s.bond_head[b], t = s.bond_tail[b];
        if (s.poly_type[h] == ptype) continue;
        double rd[3];
        for (int a = 0; a < 3; ++a) rd[a] = x[3 * t + a] - x[3 * h + a];
        const double rmag = std::sqrt(rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2]);
        for (int a = 0; a < 3; ++a)
            if (std::fabs(rd[a]) > 1.0) x[3 * t + a] = x[3 * h + a] + rd[a] / rmag;
    }

    // 2-d Maxwell velocities in the plane of the lamella
    const auto tab = maxwell_table(n);
    const std::vector<double>& table = *tab;
    const double two_pi = 2.0 * 3.141592653589793238;
    std::vector<double> vx((size_t)n), vy((size_t)n);
    double vcm[2] = {0.0, 0.0};
    for (int64_t i = 0; i < n; ++i) {
        const size_t index = (size_t)((double)table.size() * U.next());
        const double vmag = table[index];
        const double vphi = two_pi * U.next();
        vx[(size_t)i] = vmag * std::cos(vphi);
        vy[(size_t)i] = vmag * std::sin(vphi);
        vcm[0] += vx[(size_t)i];
        vcm[1] += vy[(size_t)i];
    }
    vcm[0] /= (double)n;
    vcm[1] /= (double)n;
    double vtotal = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        vx[(size_t)i] -= vcm[0];
        vy[(size_t)i] -= vcm[1];
        vtotal += vx[(size_t)i] * vx[(size_t)i] + vy[(size_t)i] * vy[(size_t)i];
    }
    vtotal = std::sqrt(vtotal / (2.0 * (double)n * d.kT));
    for (int64_t i = 0; i < n; ++i) {
        const double a = vx[(size_t)i] / vtotal, c = vy[(size_t)i] / vtotal;
        double* q = v + 3 * i;
        if (X) { q[0] = 0.0; q[1] = c; q[2] = a; }
        else if (Y) { q[0] = a; q[1] = 0.0; q[2] = c; }
        else { q[0] = a; q[1] = c; q[2] = 0.0; }
    }
    return true;
}

std::string fmt6(double v)
{
    std::ostringstream o;
    o << v;   // default ostream precision (6 significant digits), as the reference
    return o.str();
}

}  // namespace

extern "C" {

int dpd_input_parse(const char* text, dpd_input** out, char* err, int32_t errlen)
{
    if (!out || !text) return DPD_ERR_PARAM;
    *out = nullptr;
    dpd_input* in = new dpd_input();
    std::string e;
    if (!dpd::host::read_dmpci(text, in->d, e)) {
        put_err(err, errlen, e);
        delete in;
        return DPD_ERR_IO;
    }
    return finish_input(in, out, err, errlen);
}

int dpd_input_read(const char* path, dpd_input** out, char* err, int32_t errlen)
{
    if (!out || !path) return DPD_ERR_PARAM;
    *out = nullptr;
    dpd_input* in = new dpd_input();
    std::string e;
    if (!dpd::host::read_dmpci_file(path, in->d, e)) {
        put_err(err, errlen, e);
        delete in;
        return DPD_ERR_IO;
    }
    return finish_input(in, out, err, errlen);
}

void dpd_input_free(dpd_input* in) { delete in; }

int dpd_input_counts(const dpd_input* in, dpd_counts* o)
{
    if (!in || !o) return DPD_ERR_PARAM;
    const InputData& d = in->d;
    std::memset(o, 0, sizeof(*o));
    o->ntypes = d.ntypes();
    o->nbeads = (int64_t)in->s.id.size();
    o->nbonds = (int64_t)in->s.bond_head.size();
    o->npairs = (int64_t)in->s.pair_first.size();
    o->npolymers = in->s.npolymers;
    o->npolymer_types = (int32_t)d.polymers.size();
    o->nx = d.nx; o->ny = d.ny; o->nz = d.nz;
    o->wx = d.wx; o->wy = d.wy; o->wz = d.wz;
    o->density = d.density; o->kT = d.kT; o->lambda = d.lambda; o->dt = d.dt;
    o->seed = d.seed;
    o->total_time = d.total_time;
    o->sample_period = d.sample_period;
    o->restart_period = d.restart_period;
    o->ncommands = (int32_t)d.commands.size();
    o->state = d.state == "random" ? 0 : d.state == "restart" ? 1 : d.state == "lamella" ? 2 : 3;
    o->restart_state = d.restart_state;
    return DPD_OK;
}

int dpd_input_types(const dpd_input* in, char* names, int32_t len, double* a, double* gamma)
{
    if (!in) return DPD_ERR_PARAM;
    std::string all;
    for (size_t i = 0; i < in->d.bead_names.size(); ++i) all += (i ? "," : "") + in->d.bead_names[i];
    put_err(names, len, all);
    if (a) std::copy(in->a.begin(), in->a.end(), a);
    if (gamma) std::copy(in->g.begin(), in->g.end(), gamma);
    return DPD_OK;
}

int dpd_input_beads(const dpd_input* in, int32_t* id, int32_t* type, int32_t* poly_id, int32_t* poly_type)
{
    if (!in) return DPD_ERR_PARAM;
    const System& s = in->s;
    if (id) std::copy(s.id.begin(), s.id.end(), id);
    if (type) std::copy(s.type.begin(), s.type.end(), type);
    if (poly_id) std::copy(s.poly_id.begin(), s.poly_id.end(), poly_id);
    if (poly_type) std::copy(s.poly_type.begin(), s.poly_type.end(), poly_type);
    return DPD_OK;
}

int dpd_input_write_summary(const dpd_input* in, const char* path)
{
    if (!in || !path) return DPD_ERR_PARAM;
    const InputData& d = in->d;
    const System& s = in->s;
    std::ofstream o(path);
    if (!o) return DPD_ERR_IO;
    // CInitialState::SerializeAsciiFile (src/InitialState.cpp:1153-1347): counts, then the
    // bead, bond, bond-pair and polymer types in declaration order, 6 significant digits
    // (the fraction in a left-justified field of 12 at 8 digits, which stays the stream's
    // precision)
    o << "Number of polymers created\t\t" << s.npolymers << "\n";
    o << "Number of beads created\t\t\t" << s.id.size() << "\n";
    o << "Number of bonds created\t\t\t" << s.bond_head.size() << "\n";
    o << "Number of bond pairs created\t" << s.pair_first.size() << "\n";
    for (size_t t = 0; t < s.polymers_per_type.size(); ++t)
        o << "  # of type " << t << "  " << s.polymers_per_type[t] << "\n";
    const int nt = d.ntypes();
    for (int t = 0; t < nt; ++t) {
        o << "\nBead type " << t << "\n  Radius: " << d.bead_radius[t] << "\n  Cons Int: ";
        for (int j = 0; j < nt; ++j) o << d.a[t * nt + j] << " ";
        o << "\n  Diss Int: ";
        for (int j = 0; j < nt; ++j) o << d.g[t * nt + j] << " ";
        o << "\n";
    }
    for (size_t b = 0; b < d.bonds.size(); ++b)
        o << "\nBond type " << b << "\n  Spring Const/Unstr length " << d.bonds[b].k << " " << d.bonds[b].l0 << "\n";
    o << "\n";
    for (size_t b = 0; b < d.bondpairs.size(); ++b)
        o << "\nStiff bond type " << b << "\n  Modulus/Angle " << d.bondpairs[b].k << " "
          << d.bondpairs[b].phi0_deg * 3.14159265358979323846 / 180.0 << "\n";
    o << "\n";
    for (size_t t = 0; t < s.types.size(); ++t) {
        const auto& pt = s.types[t];
        o << "Polymer type " << t << "\n";
        o << "  with fraction          " << std::setw(12) << std::setprecision(8) << std::left << pt.fraction << "\n";
        o << "  containing beads       " << pt.bead_type.size() << "\n";
        o << "  and bonds              " << pt.bond_type.size() << "\n";
        o << "  Bead types\n";
        for (int bt : pt.bead_type) o << "  " << bt << " ";
        if (!pt.bond_type.empty()) {
            o << "\n  Bond types\n";
            for (int bt : pt.bond_type) o << "  " << bt << " ";
            o << "\n  Stiff bond types\n";
            for (int bt : pt.pair_type) o << "  " << bt << " ";
        }
        o << "\n\n";
    }
    return o ? DPD_OK : DPD_ERR_IO;
}

int dpd_input_bond_types(const dpd_input* in, int32_t* count, int32_t* head_type, int32_t* tail_type)
{
    if (!in) return DPD_ERR_PARAM;
    const auto& b = in->d.bonds;
    if (count) *count = (int32_t)b.size();
    for (size_t i = 0; i < b.size(); ++i) {
        if (head_type) head_type[i] = in->d.type_of(b[i].head);
        if (tail_type) tail_type[i] = in->d.type_of(b[i].tail);
    }
    return DPD_OK;
}

int dpd_input_topology(const dpd_input* in, int32_t* head, int32_t* tail, double* k, double* l0, int32_t* first,
                       int32_t* second, double* modulus, double* phi0)
{
    if (!in) return DPD_ERR_PARAM;
    const System& s = in->s;
    if (head) std::copy(s.bond_head.begin(), s.bond_head.end(), head);
    if (tail) std::copy(s.bond_tail.begin(), s.bond_tail.end(), tail);
    if (k) std::copy(s.bond_k.begin(), s.bond_k.end(), k);
    if (l0) std::copy(s.bond_l0.begin(), s.bond_l0.end(), l0);
    if (first) std::copy(s.pair_first.begin(), s.pair_first.end(), first);
    if (second) std::copy(s.pair_second.begin(), s.pair_second.end(), second);
    if (modulus) std::copy(s.pair_k.begin(), s.pair_k.end(), modulus);
    if (phi0) std::copy(s.pair_phi0.begin(), s.pair_phi0.end(), phi0);
    return DPD_OK;
}

int dpd_input_params(const dpd_input* in, int32_t device, dpd_params* p)
{
    if (!in || !p) return DPD_ERR_PARAM;
    const InputData& d = in->d;
    std::memset(p, 0, sizeof(*p));
    p->nx = d.nx; p->ny = d.ny; p->nz = d.nz;
    p->wx = d.wx; p->wy = d.wy; p->wz = d.wz;
    p->dt = d.dt;
    p->lambda = d.lambda;
    p->kT = d.kT;
    p->seed = (uint64_t)std::llabs(d.seed);   // CCNTCell::SetRNGSeed takes |seed|
    p->ntypes = d.ntypes();
    p->a = in->a.data();
    p->gamma = in->g.data();
    p->device = device;
    p->pair_kernel = DPD_PAIR_AUTO;
    return DPD_OK;
}

int dpd_input_build_state(const dpd_input* in, uint64_t seed, double* x, double* unpbc, double* v)
{
    if (!in || !x || !v) return DPD_ERR_PARAM;
    const size_t n = in->s.id.size();
    if (in->d.state != "lamella") {
        build_reference_random(in, seed, x, v);
        if (unpbc) std::copy(x, x + 3 * n, unpbc);
        return DPD_OK;
    }
    std::string err;
    if (!build_reference_lamella(in, seed, x, v, err)) return DPD_ERR_UNSUPPORTED;
    if (unpbc) std::copy(x, x + 3 * n, unpbc);
    return DPD_OK;
}

int dpd_input_command(const dpd_input* in, int32_t i, char* name, int32_t nlen, int64_t* time, char* args,
                      int32_t alen)
{
    if (!in || i < 0 || i >= (int32_t)in->d.commands.size()) return DPD_ERR_PARAM;
    const auto& c = in->d.commands[i];
    put_err(name, nlen, c.name);
    if (time) *time = c.time;
    std::string a;
    for (size_t k = 0; k < c.args.size(); ++k) a += (k ? " " : "") + c.args[k];
    put_err(args, alen, a);
    return DPD_OK;
}

int dpd_restart_read(const char* path, int64_t n, int32_t* poly_id, int32_t* poly_type, int32_t* id, int32_t* type,
                     double* x, double* unpbc, double* v, double* f)
{
    std::ifstream in(path);
    if (!in) return DPD_ERR_IO;
    int w, g, sh;
    if (!(in >> w >> g >> sh)) return DPD_ERR_IO;
    for (int64_t i = 0; i < n; ++i) {
        long pid, pt, bid, bt;
        double r, c[12];
        if (!(in >> pid >> pt >> bid >> bt >> r)) return DPD_ERR_IO;
        for (double& q : c)
            if (!(in >> q)) return DPD_ERR_IO;
        if (poly_id) poly_id[i] = (int32_t)pid;
        if (poly_type) poly_type[i] = (int32_t)pt;
        if (id) id[i] = (int32_t)bid;
        if (type) type[i] = (int32_t)bt;
        for (int a = 0; a < 3; ++a) {
            if (x) x[3 * i + a] = c[a];
            if (unpbc) unpbc[3 * i + a] = c[3 + a];
            if (v) v[3 * i + a] = c[6 + a];
            if (f) f[3 * i + a] = c[9 + a];
        }
    }
    return DPD_OK;
}

int dpd_restart_write(const char* path, int64_t n, const int32_t* poly_id, const int32_t* poly_type,
                      const int32_t* id, const int32_t* type, const double* radius_by_type, const double* L3,
                      const double* x, const double* unpbc, const double* v, const double* f)
{
    std::ofstream o(path);
    if (!o) return DPD_ERR_IO;
    o << "0 0 0\n";
    for (int64_t i = 0; i < n; ++i) {
        double xp[3];
        for (int a = 0; a < 3; ++a) {
            xp[a] = x[3 * i + a];
            if (xp[a] > L3[a]) return DPD_ERR_OUT_OF_BOX;
            if (xp[a] > L3[a] - 0.001 || xp[a] < 0.001) xp[a] = 0.001;   // CheckBeadWithinBox
        }
        o << poly_id[i] << " " << poly_type[i] << " " << id[i] << " " << type[i] << " "
          << fmt6(radius_by_type ? radius_by_type[type[i]] : 0.5) << " ";
        for (int a = 0; a < 3; ++a) o << fmt6(xp[a]) << " ";
        for (int a = 0; a < 3; ++a) o << fmt6(unpbc[3 * i + a]) << " ";
        for (int a = 0; a < 3; ++a) o << fmt6(v[3 * i + a]) << " ";
        o << fmt6(f[3 * i]) << " " << fmt6(f[3 * i + 1]) << " " << fmt6(f[3 * i + 2]) << "\n";
    }
    return o ? DPD_OK : DPD_ERR_IO;
}

int dpd_run(const char* dir, const char* run_id, int32_t device, dpd_run_stats* stats, char* err, int32_t errlen)
{
    if (!dir || !run_id) return DPD_ERR_PARAM;
    const std::string base = std::string(dir) + "/";
    dpd_input* in = nullptr;
    int rc = dpd_input_read((base + "dmpci." + run_id).c_str(), &in, err, errlen);
    if (rc) return rc;
    std::unique_ptr<dpd_input> hold(in);
    const InputData& d = in->d;
    const System& s = in->s;
    const int64_t n = (int64_t)s.id.size();
    std::vector<double> x(3 * n), u(3 * n), v(3 * n), f(3 * n, 0.0);
    if (d.state == "restart") {
        const std::string rp = base + "dmpcrs." + d.restart_run + ".con." + std::to_string(d.restart_state) + ".dat";
        std::vector<int32_t> ids(n);
        rc = dpd_restart_read(rp.c_str(), n, nullptr, nullptr, ids.data(), nullptr, x.data(), u.data(), v.data(),
                              f.data());
        if (rc) { put_err(err, errlen, "cannot read restart snapshot " + rp); return rc; }
        for (int64_t i = 0; i < n; ++i)
            if (ids[i] != s.id[i]) { put_err(err, errlen, "restart snapshot does not match the polymers"); return DPD_ERR_IO; }
    } else if (d.state == "random" || d.state == "lamella") {
        rc = dpd_input_build_state(in, (uint64_t)std::llabs(d.seed), x.data(), u.data(), v.data());
        if (rc) { put_err(err, errlen, "cannot build the starting state"); return rc; }
    } else {
        put_err(err, errlen, "State " + d.state + " is not supported by this engine");
        return DPD_ERR_UNSUPPORTED;
    }
    dpd_params p;
    dpd_input_params(in, device, &p);
    dpd_engine* e = nullptr;
    rc = dpd_create(&p, &e);
    auto fail = [&](int code) {
        put_err(err, errlen, e ? dpd_last_error(e) : "engine creation failed");
        if (e) dpd_destroy(e);
        return code;
    };
    if (rc) return fail(rc);
    if ((rc = dpd_upload_beads(e, n, s.id.data(), s.type.data(), x.data(), u.data(), v.data(), f.data(), 0)))
        return fail(rc);
    if (!s.bond_head.empty() || !s.pair_first.empty())
        if ((rc = dpd_upload_topology(e, (int64_t)s.bond_head.size(), s.bond_head.data(), s.bond_tail.data(),
                                      s.bond_k.data(), s.bond_l0.data(), (int64_t)s.pair_first.size(),
                                      s.pair_first.data(), s.pair_second.data(), s.pair_k.data(),
                                      s.pair_phi0.data())))
            return fail(rc);
    // a schedule recorded by the autotuner for this shape: the file named by DPD_TUNE_FILE,
    // else dpd_tune.json in the run's directory
    {
        const char* tf = std::getenv("DPD_TUNE_FILE");
        const std::string tune = tf ? std::string(tf) : base + "dpd_tune.json";
        std::ifstream probe(tune);
        if (probe.good() && (rc = dpd_apply_tuning(e, tune.c_str(), nullptr))) return fail(rc);
    }
    if ((rc = dpd_input_write_summary(in, (base + "dmpcis." + run_id).c_str()))) {
        put_err(err, errlen, "cannot write the dmpcis summary");
        return fail(rc);
    }
    std::ofstream hist(base + "dmpchs." + run_id);
    const double L3[3] = {d.nx * d.wx, d.ny * d.wy, d.nz * d.wz};
    // history rows (CMonitor::SaveHistoryState, src/Monitor.cpp:1750-1823): time, kT, P,
    // KE, PE, total energy (zero unless ToggleEnergyOutput is on), the mean squared
    // displacement of each bead type from its starting unwrapped position over 6 dt t
    // (src/Monitor.cpp:828-945), the mean end-to-end distance of each polymer type (head
    // and tail beads, wrapped positions with one minimum-image correction,
    // src/Monitor.cpp:707-778); each value left-justified in 12 columns, 6 digits, then a
    // space (src/TimeSeriesData.cpp:84-89)
    const std::vector<double> u0 = u;
    double dt = d.dt;   // as SetTimeStepSize leaves it
    const int ntypes = d.ntypes();
    std::vector<int64_t> per_type(ntypes, 0);
    for (int64_t i = 0; i < n; ++i) per_type[s.type[i]]++;
    std::vector<int64_t> head_of, tail_of;   // first and last bead of every polymer
    for (int64_t i = 0; i < n; ++i) {
        const int64_t p = s.poly_id[i] - 1;
        if (p >= (int64_t)head_of.size()) { head_of.resize(p + 1, -1); tail_of.resize(p + 1, -1); }
        if (head_of[p] < 0) head_of[p] = i;
        tail_of[p] = i;
    }
    bool energy_out = false;
    auto history_row = [&](int64_t t, const dpd_obs& o) {
        std::vector<double> row = {(double)t, o.kT, o.pressure, 0.0, 0.0, 0.0};
        if (energy_out) {
            double pe3[3] = {0, 0, 0};
            if (dpd_energy_output(e, 0, pe3)) return false;
            const double ke = 0.5 * (o.sum_v2[0] + o.sum_v2[1] + o.sum_v2[2]);
            row[3] = ke;
            row[4] = pe3[0] + pe3[1] + pe3[2];
            row[5] = ke + row[4];
        }
        if (dpd_download(e, x.data(), u.data(), nullptr, nullptr)) return false;
        std::vector<double> msd(ntypes, 0.0);
        for (int64_t i = 0; i < n; ++i) {
            double q = 0.0;
            for (int a = 0; a < 3; ++a) q += (u[3 * i + a] - u0[3 * i + a]) * (u[3 * i + a] - u0[3 * i + a]);
            msd[s.type[i]] += q;
        }
        for (int b = 0; b < ntypes; ++b) {
            if (per_type[b] > 0) msd[b] /= 6.0 * dt * (double)t * (double)per_type[b];
            row.push_back(msd[b]);
        }
        std::vector<double> ee(s.polymers_per_type.size(), 0.0);
        for (size_t p = 0; p < head_of.size(); ++p) {
            const int64_t hb = head_of[p], tb = tail_of[p];
            double q = 0.0;
            for (int a = 0; a < 3; ++a) {
                double dx = x[3 * hb + a] - x[3 * tb + a];
                if (dx > 0.5 * L3[a]) dx -= L3[a];
                else if (dx < -0.5 * L3[a]) dx += L3[a];
                q += dx * dx;
            }
            ee[s.poly_type[hb]] += std::sqrt(q);
        }
        for (size_t p = 0; p < ee.size(); ++p) {
            if (s.polymers_per_type[p] > 0) ee[p] /= (double)s.polymers_per_type[p];
            row.push_back(ee[p]);
        }
        for (double val : row) hist << std::setw(12) << std::setprecision(6) << std::left << val << " ";
        hist << "\n";
        return true;
    };
    dpd_run_stats st{};
    st.nbeads = n;
    size_t next_cmd = 0;
    bool thermo_on = true, cons_on = true, forces_on = true;
    int64_t renorm_period = 0;   // RenormaliseMomenta: 0 = off
    // bond (pair) type of a command's name: the declared names joined, read from either end
    // (CInputData::AddBondType / AddBondPairType, src/InputData.cpp:556-598, :626-674)
    auto bond_type_of = [&](const std::string& name) -> int {
        for (size_t b = 0; b < d.bonds.size(); ++b)
            if (name == d.bonds[b].head + d.bonds[b].tail || name == d.bonds[b].tail + d.bonds[b].head) return (int)b;
        return -1;
    };
    auto pair_type_of = [&](const std::string& name) -> int {
        for (size_t b = 0; b < d.bondpairs.size(); ++b) {
            const auto& q = d.bondpairs[b];
            if (name == q.n1 + q.n2 + q.n3 || name == q.n3 + q.n2 + q.n1) return (int)b;
        }
        return -1;
    };
    // bond parameters as the commands leave them
    std::vector<double> bk = s.bond_k, bl0 = s.bond_l0, pk = s.pair_k, pphi = s.pair_phi0;
    auto upload_topology = [&]() {
        return dpd_upload_topology(e, (int64_t)s.bond_head.size(), s.bond_head.data(), s.bond_tail.data(),
                                   bk.data(), bl0.data(), (int64_t)s.pair_first.size(), s.pair_first.data(),
                                   s.pair_second.data(), pk.data(), pphi.data());
    };
    std::vector<double> a = in->a, g = in->g;
    double seconds = 0.0;
    int64_t t = 0;
    while (t < d.total_time) {
        // commands due before step t+1 (CSimBox::Run executes them before Evolve)
        while (next_cmd < d.commands.size() && d.commands[next_cmd].time <= t + 1) {
            const auto& c = d.commands[next_cmd++];
            if (c.time < t + 1) continue;
            if (c.name == "ToggleDPDBeadThermostat") { thermo_on = !thermo_on; rc = dpd_toggle(e, DPD_TOGGLE_THERMOSTAT, thermo_on); }
            else if (c.name == "ToggleDPDBeadConservativeForces") { cons_on = !cons_on; rc = dpd_toggle(e, DPD_TOGGLE_CONSERVATIVE, cons_on); }
            else if (c.name == "ToggleDPDBeadForces") { forces_on = !forces_on; rc = dpd_toggle(e, DPD_TOGGLE_FORCES, forces_on); }
            else if (c.name == "SetTimeStepSize" && !c.args.empty()) { dt = std::atof(c.args[0].c_str()); rc = dpd_set_timestep(e, dt); }
            else if ((c.name == "SetDPDBeadConsInt" || c.name == "SetDPDBeadDissInt" ||
                      c.name == "SetDPDBeadConsIntByType" || c.name == "SetDPDBeadDissIntByType") && c.args.size() >= 3) {
                // by bead names, or by type indices for the ByType forms (src/SimBox.cpp:3227-3446)
                const bool by_type = c.name.size() > 6 && c.name.compare(c.name.size() - 6, 6, "ByType") == 0;
                const int i = by_type ? std::atoi(c.args[0].c_str()) : d.type_of(c.args[0]);
                const int j = by_type ? std::atoi(c.args[1].c_str()) : d.type_of(c.args[1]);
                if (i < 0 || j < 0 || i >= d.ntypes() || j >= d.ntypes()) continue;   // the reference logs a failed command
                std::vector<double>& m = c.name.compare(0, 17, "SetDPDBeadConsInt") == 0 ? a : g;
                m[i * d.ntypes() + j] = m[j * d.ntypes() + i] = std::atof(c.args[2].c_str());
                rc = dpd_set_interactions(e, a.data(), g.data());
            } else if (c.name == "SetBondStrength" && c.args.size() >= 3) {
                // CSimBox::SetBondStrength (src/SimBox.cpp:3505-3545): k and l0 of every bond of the type
                const int bt = bond_type_of(c.args[0]);
                if (bt < 0) continue;
                const double k = std::atof(c.args[1].c_str()), l0 = std::atof(c.args[2].c_str());
                for (size_t q = 0; q < s.bond_type.size(); ++q)
                    if (s.bond_type[q] == bt) { bk[q] = k; bl0[q] = l0; }
                rc = upload_topology();
            } else if (c.name == "SetBondStiffness" && c.args.size() >= 3) {
                // CSimBox::SetBondStiffness (src/SimBox.cpp:3447-3503); the angle is read in degrees
                const int pt = pair_type_of(c.args[0]);
                if (pt < 0) continue;
                const double k = std::atof(c.args[1].c_str());
                const double phi0 = std::atof(c.args[2].c_str()) * 3.141592653589793238 / 180.0;
                for (size_t q = 0; q < s.pair_type.size(); ++q)
                    if (s.pair_type[q] == pt) { pk[q] = k; pphi[q] = phi0; }
                rc = upload_topology();
            } else if (c.name == "ToggleEnergyOutput") {
                // CMonitor::InternalToggleEnergyOutput (src/Monitor.cpp:3171-3177)
                energy_out = !energy_out;
                rc = DPD_OK;
            } else if (c.name == "ZeroTotalMom") {
                rc = dpd_zero_total_momentum(e, -1);
            } else if (c.name == "RenormaliseMomenta" && c.args.size() >= 2) {
                // ccRenormaliseMomenta: Period, flag (src/ccRenormaliseMomenta.cpp:107)
                renorm_period = std::atoi(c.args[1].c_str()) ? std::max<int64_t>(1, std::atoll(c.args[0].c_str())) : 0;
            } else {
                continue;   // commands outside the step engine are not executed
            }
            if (rc) return fail(rc);
            st.commands_applied++;
        }
        // CSimBox::Run renormalises before Evolve when m_SimTime % Period == 0
        // (src/SimBox.cpp:1313-1317); m_SimTime of the coming step is t + 1
        if (renorm_period > 0 && (t + 1) % renorm_period == 0)
            if ((rc = dpd_renormalise_momenta(e, -1))) return fail(rc);
        int64_t stop = d.total_time;
        if (next_cmd < d.commands.size()) stop = std::min(stop, d.commands[next_cmd].time - 1);
        if (renorm_period > 0) stop = std::min(stop, ((t + 1) / renorm_period + 1) * renorm_period - 1);
        stop = std::min(stop, (t / d.sample_period + 1) * d.sample_period);
        stop = std::min(stop, (t / d.restart_period + 1) * d.restart_period);
        if (stop <= t) stop = t + 1;
        const bool sample = stop % d.sample_period == 0;
        const auto t0 = std::chrono::steady_clock::now();
        // captured step graphs are replayed for repeated call lengths
        if ((rc = dpd_step(e, stop - t, DPD_STEP_GRAPH | (sample ? DPD_STEP_OBSERVE : 0)))) return fail(rc);
        seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        t = stop;
        if (sample) {
            dpd_obs o;
            dpd_observables(e, &o);
            if (!history_row(t, o)) return fail(DPD_ERR_HIP);
            st.last_kT = o.kT;
            st.last_pressure = o.pressure;
            st.samples++;
        }
        if (t % d.restart_period == 0) {
            if ((rc = dpd_download(e, x.data(), u.data(), v.data(), f.data()))) return fail(rc);
            const std::string rp = base + "dmpcrs." + std::string(run_id) + ".con." + std::to_string(t) + ".dat";
            if ((rc = dpd_restart_write(rp.c_str(), n, s.poly_id.data(), s.poly_type.data(), s.id.data(),
                                        s.type.data(), d.bead_radius.data(), L3, x.data(), u.data(), v.data(),
                                        f.data())))
                return fail(rc);
        }
    }
    st.steps = t;
    st.seconds = seconds;
    st.bead_steps_per_s = seconds > 0 ? (double)n * (double)t / seconds : 0.0;
    if (stats) *stats = st;
    dpd_destroy(e);
    return DPD_OK;
}

}  // extern "C"
