// Output writers of the engine: global statistics, fluid probes, particle
// CSV, VTK (XML .vti/.vtp with .pvd collections, or legacy .vtk), in the
// reference's file layouts (io/global_stats_writer.cpp, fluid_probe_writer.cpp,
// csv_writer.cpp, vtk_writer.cpp; configuration io/output_config.h).
//
// The device keeps stepping while files are written: at an output step the
// engine enqueues one snapshot of what the due writers need (sf_snapshot_async:
// device reductions for the statistics, a gather for the probes, fields and
// nodes only when VTK or CSV output is due) into pinned host buffers, and a
// writer thread formats and writes once the copies have landed.  Large arrays
// are formatted by several threads, each into its own buffer, written in order.
#pragma once
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/softflow.h"

namespace softflow {

class Capsule;
struct SimulationParams;

// io/output_config.h: field toggles, filter, probes and the master config
struct FluidOutputFields {
    bool density = true;
    bool velocity = true;
    bool pressure = true;
    bool vorticity = false;
    bool strain_rate = false;
    bool ibm_force = false;
    bool node_type = true;
    bool velocity_mag = true;
    bool component_density = true;
    bool concentration = true;
};

struct ParticleVTKFields {
    bool velocity = true;
    bool force = true;
    bool particle_id = true;
    bool particle_type = true;
    bool particle_group = true;
    bool local_strain = false;
    bool local_curvature = false;
};

struct ParticleOutputFields {
    bool position = true;
    bool velocity = true;
    bool force = false;
    bool diameter = false;
    bool type = true;
    bool group = false;
    bool deformation = false;
    bool area_volume = false;
    bool orientation = false;
    bool angular_vel = false;
    bool com_position = true;
    bool com_velocity = true;
};

struct ParticleFilter {
    enum FilterType { ALL, BY_ID, BY_TYPE, BY_GROUP, BY_REGION, COMBINED };
    FilterType filter_type = ALL;
    std::vector<int> selected_ids, selected_types, selected_groups;
    double region_xmin = -1e30, region_xmax = 1e30;
    double region_ymin = -1e30, region_ymax = 1e30;
    bool passes(int id, int type_id, int group, double cx, double cy) const;
};

struct FluidProbe {
    int i = 0, j = 0;
    std::string label;
};

struct ExtraCSVConfig {
    std::string filename;
    int dump_every = 100;
    ParticleOutputFields fields;
    ParticleFilter filter;
    bool append = true;
};

struct OutputConfig {
    std::string output_dir = "output";
    int vtk_dump_every = 1000;
    std::string vtk_format = "ascii";
    int csv_dump_every = 100;
    std::string csv_format = "csv";
    bool csv_append = true;
    int probe_dump_every = 10;
    int stats_dump_every = 100;
    FluidOutputFields fluid_fields;
    ParticleVTKFields particle_vtk_fields;
    ParticleOutputFields csv_fields;
    ParticleFilter csv_filter;
    std::vector<ExtraCSVConfig> extra_csv;
    std::vector<FluidProbe> probes;
};

// One output step's data on the host.
struct Frame {
    int step = 0;
    double time = 0.0;
    bool vtk = false, csv = false, probes = false, stats = false;
    int nx = 0, ny = 0, nsp = 0;
    // pinned buffers: N-sized fields, nodes, probes, statistics
    double *rho = nullptr, *ux = nullptr, *uy = nullptr, *Fx = nullptr, *Fy = nullptr;
    double* C[4] = {};
    double *x = nullptr, *y = nullptr, *vx = nullptr, *vy = nullptr, *fx = nullptr, *fy = nullptr;
    double* probe_vals = nullptr;
    double stat_vals[12] = {};
    std::vector<int> flags;                  // cell types at the step
    std::vector<Capsule> caps;               // capsule metadata; nodes filled from the snapshot
    int64_t ticket = 0;
    int slot = -1;                           // pinned buffer set
};

// A CSV particle writer's state (io/csv_writer.h).
struct CsvState {
    std::string filename, format;
    bool append = true;
    ParticleOutputFields fields;
    ParticleFilter filter;
    FILE* fp = nullptr;
    bool header = false;
};

class OutputManager {
public:
    OutputManager(const SimulationParams& p, int nsp);
    ~OutputManager();
    OutputManager(const OutputManager&) = delete;
    OutputManager& operator=(const OutputManager&) = delete;

    // the device context snapshots are taken from (set at initialize())
    void attach(sf_ctx* ctx);
    // intervals in effect (simulation.cpp:737-766): 0 = off
    int vtkInterval() const;
    int csvInterval() const;
    bool due(int64_t step) const;
    // the first step index >= from at which some writer is due (-1: none)
    int64_t nextDue(int64_t from) const;
    // enqueue the snapshot of step `step` and hand it to the writer thread
    void capture(int64_t step, const std::vector<Capsule>& caps, const std::vector<int>& flags);
    // wait until every captured step is written
    void drain();
    // Simulation::finalize (simulation.cpp:801-812): drain, .pvd files, close
    void finalize();

    // configuration (simulation.cpp:971-1027)
    void setOutputConfig(const OutputConfig& c);
    void setVTKOutput(int dump_every, const std::string& format, const FluidOutputFields& ff,
                      const ParticleVTKFields& pf);
    void setCSVOutput(int dump_every, const std::string& format, const ParticleOutputFields& fields,
                      const ParticleFilter& filter, bool append);
    void addExtraCSVOutput(const std::string& filename, const ParticleOutputFields& fields,
                           const ParticleFilter& filter);
    void addFluidProbe(int i, int j, const std::string& label);
    const OutputConfig& config() const { return cfg_; }

private:
    void worker();
    void write(Frame& f);
    void writeStats(const Frame& f);
    void writeProbes(const Frame& f);
    void writeCsv(CsvState& w, const Frame& f);
    void writeFluidVtk(const Frame& f);
    void writeCapsulesVtk(const Frame& f);
    void writePvd();
    void ensureDir(const std::string& d);
    void freeSlots();
    void uploadProbes();

    sf_ctx* ctx_;
    std::string dir_;
    int nx_, ny_, nsp_, periodic_nx_;
    double dt_, rho0_;
    int vtk_every_, csv_every_, probe_every_, stats_every_, output_interval_;
    bool legacy_ = false;
    bool has_scalar_;
    OutputConfig cfg_;
    CsvState csv_;
    std::vector<std::unique_ptr<CsvState>> extra_csv_;
    FILE* stats_fp_ = nullptr;
    bool stats_header_ = false;
    FILE* probe_fp_ = nullptr;
    bool probe_header_ = false;
    bool probes_dirty_ = true;
    std::vector<int> recorded_steps_;
    std::vector<double> recorded_times_;
    std::vector<std::string> made_dirs_;
    // pinned buffer sets, two so that one can be written while the next fills
    struct Slot {
        double* fields = nullptr;     // 5N + nsp N
        double* nodes = nullptr;      // 6K
        double* probes = nullptr;     // 3 nprobes
        size_t nfields = 0, nnodes = 0, nprobes = 0;
        bool busy = false;
    };
    Slot slots_[2];
    // writer thread
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::unique_ptr<Frame>> queue_;
    int inflight_ = 0;
    bool stop_ = false;
    std::string error_;
    std::thread thread_;
};

}  // namespace softflow
