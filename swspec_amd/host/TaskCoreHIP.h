// TaskCoreHIP.h -- the MI355X compute core as a TaskCore plugin.
//
// Implements the reference's abstract core interface, class TaskCore
// (src/TaskCore.h:36-89), on top of the C ABI in include/swspec_hip.h.  It is
// selected by the dispatcher's compile-time switch (src/TaskDispatcher.h:28-41)
// exactly like TaskCoreIPP; nothing else in the reference changes.  Built
// inside the reference tree (include path = its src/ directory).
#ifndef TASKCOREHIP_H
#define TASKCOREHIP_H

#include "TaskCore.h"
#include "Settings.h"
#include "Buffer.h"
#include "swspec_hip.h"

#include <ostream>

class TaskCoreHIP : public TaskCore
{
public:
    explicit TaskCoreHIP(int mrank);
    ~TaskCoreHIP();

    // TaskCore::prepare: decoder, window, FFT plan, phase-cal rotator and
    // device buffers for this core (TaskCoreIPP::prepare, TaskCoreIPP.cpp:70-145)
    void prepare(swspect_settings_t* settings);

    // TaskCore::run: asynchronous; returns 0 when the work was queued
    int run(Buffer** inbuf, Buffer** outbuf, Buffer** xpolbuf, Buffer** pcalbuf);

    // TaskCore::join: number of spectra in the output buffers, -1 on failure
    int join();

    // TaskCore::finalize: release device resources, log the calculation time
    int finalize();

    // TaskCore::combineResults: add this core's results into the dispatcher's
    // combined buffers and clear them (TaskCoreIPP.cpp:311-349)
    int combineResults(Buffer** outbuf, Buffer** pcalbuf);

    // TaskCore::resetBuffer: zero a buffer's floats
    void resetBuffer(Buffer* buf);

private:
    int rank;
    swspec_hip_core* core;
    swspect_settings_t* cfg;
    Buffer** buf_out;
    Buffer** bufxpol_out;
    Buffer** bufpcal_out;
    double total_runtime;
    long long total_spectra;
    bool failed;
};

#endif // TASKCOREHIP_H
