// simbox_engine.cpp -- CSimBoxEngine: the reference's state uploaded to the MI355X engine
// (include/dpd_engine.h) and stepped there in place of CSimBox::Evolve
// (src/SimBox.cpp:977-1120).  Built only against the reference's headers.
#include "simbox_engine.h"

#include <cmath>
#include <cstdlib>
#include <iostream>

#include "SimDefs.h"
#include "ISimBox.h"
#include "AbstractBead.h"
#include "Bond.h"
#include "BondPair.h"

static CSimBoxEngine* g_pEngine = 0;

CSimBoxEngine& CSimBoxEngine::Get(const ISimBox& box, bool unsupported, long simTime)
{
    // Walls and body forces move no bead in the engine's step: refuse them rather than run a
    // different simulation (CSimBox::Evolve's wall and gravity branches, src/SimBox.cpp:1040-1100).
    if (unsupported) {
        std::cerr << "MI355X engine: walls and gravity are not supported" << std::endl;
        std::exit(1);
    }
    if (!g_pEngine)
        g_pEngine = new CSimBoxEngine(box, simTime - 1);
    return *g_pEngine;
}

bool CSimBoxEngine::Synced()
{
    return g_pEngine && g_pEngine->m_bSynced;
}

void CSimBoxEngine::Check(int rc, const char* what) const
{
    if (rc == DPD_OK)
        return;
    std::cerr << "MI355X engine: " << what << " failed: "
              << (m_pEngine ? dpd_last_error(m_pEngine) : "no engine") << std::endl;
    std::exit(1);
}

CSimBoxEngine::CSimBoxEngine(const ISimBox& s, long step0) : m_pEngine(0), m_bSynced(false)
{
    // interactions: the symmetric a_ij and gamma_ij the CCNTCell statics hold
    const long nt = s.GetBeadTypeTotal();
    std::vector<double> a(nt * nt), g(nt * nt);
    for (long i = 0; i < nt; ++i)
        for (long j = 0; j < nt; ++j) {
            a[i * nt + j] = s.GetDPDBeadConsInt(i, j);
            g[i * nt + j] = s.GetDPDBeadDissInt(i, j);
        }

    dpd_params p = {};
    p.nx = (int32_t)s.GetCNTXCellNo();
    p.ny = (int32_t)s.GetCNTYCellNo();
    p.nz = (int32_t)s.GetCNTZCellNo();
    p.wx = s.GetCNTXCellWidth();
    p.wy = s.GetCNTYCellWidth();
    p.wz = s.GetCNTZCellWidth();
    p.dt = s.GetStepSize();
    p.lambda = s.GetLambda();
    p.kT = s.GetTemperature();
    p.seed = (uint64_t)std::labs(s.GetRNGSeed());     // CCNTCell::SetRNGSeed keeps |seed|
    p.ntypes = (int32_t)nt;
    p.a = a.data();
    p.gamma = g.data();
    p.device = 0;
    p.pair_kernel = DPD_PAIR_AUTO;
    Check(dpd_create(&p, &m_pEngine), "dpd_create");

    // beads in the order of the CNT cells at the first step, FP64 xyz-interleaved
    const AbstractBeadVector beads = s.GetBeads();
    const size_t n = beads.size();
    m_vBeads.assign(beads.begin(), beads.end());
    std::unordered_map<const CAbstractBead*, int32_t> index;
    std::vector<int32_t> id(n), type(n);
    m_x.resize(3 * n); m_u.resize(3 * n); m_v.resize(3 * n); m_f.resize(3 * n);
    for (size_t i = 0; i < n; ++i) {
        const CAbstractBead* b = beads[i];
        index[b] = (int32_t)i;
        id[i] = (int32_t)b->GetId();
        type[i] = (int32_t)b->GetType();
        const double x[3] = {b->GetXPos(), b->GetYPos(), b->GetZPos()};
        const double u[3] = {b->GetunPBCXPos(), b->GetunPBCYPos(), b->GetunPBCZPos()};
        const double v[3] = {b->GetXMom(), b->GetYMom(), b->GetZMom()};
        const double f[3] = {b->GetXForce(), b->GetYForce(), b->GetZForce()};
        for (int c = 0; c < 3; ++c) {
            m_x[3 * i + c] = x[c]; m_u[3 * i + c] = u[c];
            m_v[3 * i + c] = v[c]; m_f[3 * i + c] = f[c];
        }
    }
    Check(dpd_upload_beads(m_pEngine, (int64_t)n, id.data(), type.data(), m_x.data(), m_u.data(),
                           m_v.data(), m_f.data(), (int64_t)step0), "dpd_upload_beads");

    // bonds by bead index, bond pairs by bond index
    const BondVector& bonds = s.GetBonds();
    std::unordered_map<const CBond*, int32_t> bindex;
    std::vector<int32_t> head(bonds.size()), tail(bonds.size());
    std::vector<double> k(bonds.size()), l0(bonds.size());
    for (size_t i = 0; i < bonds.size(); ++i) {
        const CBond* bd = bonds[i];
        bindex[bd] = (int32_t)i;
        head[i] = index.at(bd->GetHead());
        tail[i] = index.at(bd->GetTail());
        k[i] = bd->GetSprConst();
        l0[i] = bd->GetUnStrLength();
    }
    const BondPairVector& pairs = s.GetBondPairs();
    std::vector<int32_t> first(pairs.size()), second(pairs.size());
    std::vector<double> modulus(pairs.size()), phi0(pairs.size());
    for (size_t i = 0; i < pairs.size(); ++i) {
        const CBondPair* bp = pairs[i];
        first[i] = bindex.at(bp->GetFirst());
        second[i] = bindex.at(bp->GetSecond());
        modulus[i] = bp->GetModulus();
        phi0[i] = bp->GetPhi0();
    }
    Check(dpd_upload_topology(m_pEngine, (int64_t)bonds.size(), head.data(), tail.data(), k.data(),
                              l0.data(), (int64_t)pairs.size(), first.data(), second.data(),
                              modulus.data(), phi0.data()), "dpd_upload_topology");
}

CSimBoxEngine::~CSimBoxEngine()
{
    dpd_destroy(m_pEngine);
}

void CSimBoxEngine::Step(bool sync)
{
    Check(dpd_step(m_pEngine, 1, sync ? DPD_STEP_OBSERVE : 0u), "dpd_step");
    m_bSynced = sync;
    if (sync)
        CopyToBeads();
}

void CSimBoxEngine::RenormaliseMomenta()
{
    Check(dpd_renormalise_momenta(m_pEngine, -1), "dpd_renormalise_momenta");
}

void CSimBoxEngine::CopyToBeads()
{
    Check(dpd_download(m_pEngine, m_x.data(), m_u.data(), m_v.data(), m_f.data()), "dpd_download");
    for (size_t i = 0; i < m_vBeads.size(); ++i) {
        CAbstractBead* b = m_vBeads[i];
        const double* x = &m_x[3 * i];
        const double* u = &m_u[3 * i];
        const double* v = &m_v[3 * i];
        const double* f = &m_f[3 * i];
        b->SetXPos(x[0]); b->SetYPos(x[1]); b->SetZPos(x[2]);
        b->SetunPBCXPos(u[0]); b->SetunPBCYPos(u[1]); b->SetunPBCZPos(u[2]);
        b->SetXMom(v[0]); b->SetYMom(v[1]); b->SetZMom(v[2]);
        b->SetXForce(f[0]); b->SetYForce(f[1]); b->SetZForce(f[2]);
    }
}
