// simbox_engine.h -- the engine behind CSimBox::Run's step loop.
//
// Compiled inside the reference's source tree (its headers on the include path).  The only
// edits to the reference are three lines of CSimBox::Run (src/SimBox.cpp:1316, :1320, :1322):
//
//     UpdateRenormalisedMom();  ->  DPD_SIMBOX_ENGINE.RenormaliseMomenta();
//     Evolve();                 ->  DPD_SIMBOX_ENGINE.Step(TimeToSample() || TimeToDisplay()
//                                                            || TimeToRestart());
//     CNTCellCheck();           ->  if (CSimBoxEngine::Synced()) CNTCellCheck();
//
// and an #include of this header.  The class layout of CSimBox does not change, so the
// reference's other objects link unchanged.  On the steps that sample, display or save a
// restart, Step() copies positions, unwrapped positions, momenta and forces back into the
// CBead objects before CSimBox::Sample / SaveCurrentState / SaveRestartState read them, and
// CNTCellCheck re-files the beads into their CNT cells; on the other steps the beads stay in
// device memory only.
#pragma once

#include <unordered_map>
#include <vector>

#include "dpd_engine.h"

// The engine of the CSimBox whose member function names it (CSimBox::Run): its ISimBox, its
// wall and gravity switches and its current time.
#define DPD_SIMBOX_ENGINE CSimBoxEngine::Get(*GetISimBox(), IsWallOn() || IsGravityOn(), m_SimTime)

class ISimBox;
class CAbstractBead;

class CSimBoxEngine {
public:
    // The engine of this process, created on the first call from the state of the box behind
    // `box` (beads, bonds, bond pairs, CNT cells, interactions through ISimBox's accessors),
    // whose simulation time is simTime - 1.  A failure to create it stops the process with
    // dpd_last_error's message, as CSimBox's constructor does for an invalid state
    // (src/SimBox.cpp:436-445).  Walls and body forces are refused (`unsupported`).
    static CSimBoxEngine& Get(const ISimBox& box, bool unsupported, long simTime);
    // Whether the last Step() copied the state back into the beads.
    static bool Synced();

    void Step(bool sync);            // one CSimBox::Evolve
    void RenormaliseMomenta();       // CSimBox::UpdateRenormalisedMom
    void CopyToBeads();

    ~CSimBoxEngine();

private:
    explicit CSimBoxEngine(const ISimBox& box, long step0);
    void Check(int rc, const char* what) const;

    dpd_engine* m_pEngine;
    std::vector<CAbstractBead*> m_vBeads;    // upload order
    std::vector<double> m_x, m_u, m_v, m_f;
    bool m_bSynced;
};
