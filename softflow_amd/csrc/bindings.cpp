// pybind11 module softflow_core: the reference's Python surface
// (python/bindings.cpp) over the MI355X engine.
//
// Field views are zero-copy, writable numpy arrays over the host mirror.
// While a view is alive it holds a lease on its array: the mirror is refreshed
// after every step, so the view shows the current state, and writes into it
// reach the device before the next step (LatticeField::leaseView).
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "engine.h"

namespace py = pybind11;
using namespace softflow;

namespace {

// The lease a view holds on one of the field's arrays; released when the
// last array derived from the view is gone.

// The lease holds a strong reference to the Python object of the field (and
// through it the Simulation), so the C++ object outlives every view: the
// release runs before that reference is dropped, whatever order the
// interpreter tears objects down in.
template <class Owner>
py::capsule lease(py::handle self, int kind, Real** data) {
    Owner& f = self.cast<Owner&>();
    *data = f.leaseView(kind);
    if (!*data) throw py::index_error("no such array");
    struct Lease {
        Owner* owner;
        PyObject* pyowner;
        int kind;
    };
    self.inc_ref();
    return py::capsule(new Lease{&f, self.ptr(), kind}, [](void* p) {
        auto* l = static_cast<Lease*>(p);
        l->owner->releaseView(l->kind);
        Py_DECREF(l->pyowner);
        delete l;
    });
}

template <class Owner>
py::array view2d(py::handle self, int kind) {
    Owner& f = self.cast<Owner&>();
    Real* data = nullptr;
    py::capsule owner = lease<Owner>(self, kind, &data);
    const int ny = f.getNy(), nx = f.getNx();
    return py::array(py::dtype::of<Real>(), {ny, nx},
                     {static_cast<py::ssize_t>(nx) * static_cast<py::ssize_t>(sizeof(Real)),
                      static_cast<py::ssize_t>(sizeof(Real))},
                     data, owner);
}

// LBMSolver handle exposing field() as the reference does (bindings.cpp:456-475)
struct LBMSolverView {
    Simulation* sim;
    LatticeField& field() { return sim->field(); }
    int getNx() const { return sim->params().nx; }
    int getNy() const { return sim->params().ny; }
    int getStep() const { return sim->currentStep(); }
    bool isPeriodicX() const { return sim->params().fluid.boundary_type == BoundaryType::PERIODIC; }
    bool isPeriodicY() const { return sim->params().fluid.periodic_y; }
};

}  // namespace

PYBIND11_MODULE(softflow_core, m) {
    m.doc() = "SoftFlow per-timestep engine on MI355X (gfx950) HIP kernels";

    py::class_<Vec2d>(m, "Vec2d")
        .def(py::init<>())
        .def(py::init<Real, Real>(), py::arg("x"), py::arg("y"))
        .def_readwrite("x", &Vec2d::x)
        .def_readwrite("y", &Vec2d::y)
        .def("norm", &Vec2d::norm)
        .def("__repr__", [](const Vec2d& v) {
            return "Vec2d(" + std::to_string(v.x) + ", " + std::to_string(v.y) + ")";
        });

    py::enum_<CellType>(m, "CellType")
        .value("FLUID", CellType::FLUID).value("SOLID", CellType::SOLID)
        .value("INLET", CellType::INLET).value("OUTLET", CellType::OUTLET)
        .value("EMPTY", CellType::EMPTY);
    py::enum_<BoundaryType>(m, "BoundaryType")
        .value("INLET_OUTLET", BoundaryType::INLET_OUTLET)
        .value("PERIODIC", BoundaryType::PERIODIC)
        .value("CLOSED", BoundaryType::CLOSED);
    py::enum_<CollisionModel>(m, "CollisionModel")
        .value("BGK", CollisionModel::BGK).value("MRT", CollisionModel::MRT)
        .value("REGULARIZED", CollisionModel::REGULARIZED);
    py::enum_<MembraneModel>(m, "MembraneModel")
        .value("HOOKEAN", MembraneModel::HOOKEAN).value("NEO_HOOKEAN", MembraneModel::NEO_HOOKEAN)
        .value("SKALAK", MembraneModel::SKALAK).value("WLC", MembraneModel::WLC);
    py::enum_<CapsuleShape>(m, "CapsuleShape")
        .value("CIRCLE", CapsuleShape::CIRCLE).value("ELLIPSE", CapsuleShape::ELLIPSE)
        .value("BICONCAVE", CapsuleShape::BICONCAVE).value("FIBER", CapsuleShape::FIBER);
    py::enum_<Precision>(m, "Precision")
        .value("FP64", Precision::FP64).value("FP32", Precision::FP32);

    py::class_<FluidParams>(m, "FluidParams")
        .def(py::init<>())
        .def_readwrite("tau", &FluidParams::tau)
        .def_readwrite("rho0", &FluidParams::rho0)
        .def_readwrite("inlet_velocity", &FluidParams::inlet_velocity)
        .def_readwrite("outlet_density", &FluidParams::outlet_density)
        .def_readwrite("boundary_type", &FluidParams::boundary_type)
        .def_readwrite("body_force_x", &FluidParams::body_force_x)
        .def_readwrite("body_force_y", &FluidParams::body_force_y)
        .def_readwrite("gravity_x", &FluidParams::gravity_x)
        .def_readwrite("gravity_y", &FluidParams::gravity_y)
        .def_readwrite("apply_gravity_to_fluid", &FluidParams::apply_gravity_to_fluid)
        .def_readwrite("apply_gravity_to_capsules", &FluidParams::apply_gravity_to_capsules)
        .def_readwrite("collision_model", &FluidParams::collision_model)
        .def_readwrite("use_mrt", &FluidParams::use_mrt)
        .def_readwrite("mrt_se", &FluidParams::mrt_se)
        .def_readwrite("mrt_s_eps", &FluidParams::mrt_s_eps)
        .def_readwrite("mrt_sq", &FluidParams::mrt_sq)
        .def_readwrite("periodic_y", &FluidParams::periodic_y)
        .def_readwrite("capsule_periodic_x", &FluidParams::capsule_periodic_x)
        .def_readwrite("top_wall_velocity", &FluidParams::top_wall_velocity)
        .def_readwrite("bottom_wall_velocity", &FluidParams::bottom_wall_velocity)
        .def_readwrite("use_interpolated_bb", &FluidParams::use_interpolated_bb)
        .def_readwrite("viscosity_contrast", &FluidParams::viscosity_contrast)
        .def_readwrite("viscosity_update_interval", &FluidParams::viscosity_update_interval);

    py::class_<MembraneParams>(m, "MembraneParams")
        .def(py::init<>())
        .def_readwrite("model", &MembraneParams::model)
        .def_readwrite("k_stretch", &MembraneParams::k_stretch)
        .def_readwrite("G_s", &MembraneParams::G_s)
        .def_readwrite("C_skalak", &MembraneParams::C_skalak)
        .def_readwrite("wlc_L_max_ratio", &MembraneParams::wlc_L_max_ratio)
        .def_readwrite("wlc_kBT_p", &MembraneParams::wlc_kBT_p)
        .def_readwrite("wlc_k_pow", &MembraneParams::wlc_k_pow)
        .def_readwrite("k_bend", &MembraneParams::k_bend)
        .def_readwrite("use_helfrich_bending", &MembraneParams::use_helfrich_bending)
        .def_readwrite("kappa_0", &MembraneParams::kappa_0)
        .def_readwrite("k_area", &MembraneParams::k_area)
        .def_readwrite("k_perimeter", &MembraneParams::k_perimeter)
        .def_readwrite("gamma_visc", &MembraneParams::gamma_visc)
        .def_readwrite("eta_membrane", &MembraneParams::eta_membrane)
        .def_readwrite("viscosity_ratio", &MembraneParams::viscosity_ratio)
        .def_readwrite("density", &MembraneParams::density)
        .def_readwrite("shape", &MembraneParams::shape)
        .def_readwrite("aspect_ratio", &MembraneParams::aspect_ratio)
        .def_readwrite("indent_depth", &MembraneParams::indent_depth)
        .def_readwrite("is_rigid", &MembraneParams::is_rigid);

    py::class_<RepulsionParams>(m, "RepulsionParams")
        .def(py::init<>())
        .def_readwrite("epsilon", &RepulsionParams::epsilon)
        .def_readwrite("sigma", &RepulsionParams::sigma)
        .def_readwrite("r_cut", &RepulsionParams::r_cut)
        .def_readwrite("power", &RepulsionParams::power)
        .def_readwrite("damping_normal", &RepulsionParams::damping_normal)
        .def_readwrite("friction_coeff", &RepulsionParams::friction_coeff);

    py::class_<LubricationParams>(m, "LubricationParams")
        .def(py::init<>())
        .def_readwrite("enabled", &LubricationParams::enabled)
        .def_readwrite("h_threshold", &LubricationParams::h_threshold)
        .def_readwrite("h_min", &LubricationParams::h_min);
    py::class_<AdhesionParams>(m, "AdhesionParams")
        .def(py::init<>())
        .def_readwrite("enabled", &AdhesionParams::enabled)
        .def_readwrite("k_on", &AdhesionParams::k_on)
        .def_readwrite("k_off", &AdhesionParams::k_off)
        .def_readwrite("k_bond", &AdhesionParams::k_bond)
        .def_readwrite("d_bond", &AdhesionParams::d_bond)
        .def_readwrite("F_crit", &AdhesionParams::F_crit)
        .def_readwrite("max_bonds_per_node", &AdhesionParams::max_bonds_per_node)
        .def_readwrite("use_catch_slip", &AdhesionParams::use_catch_slip)
        .def_readwrite("k_off_catch", &AdhesionParams::k_off_catch)
        .def_readwrite("F_catch", &AdhesionParams::F_catch)
        .def_readwrite("k_off_slip", &AdhesionParams::k_off_slip)
        .def_readwrite("F_slip", &AdhesionParams::F_slip)
        .def_readwrite("wall_adhesion", &AdhesionParams::wall_adhesion)
        .def_readwrite("wall_k_on", &AdhesionParams::wall_k_on)
        .def_readwrite("wall_k_off", &AdhesionParams::wall_k_off)
        .def_readwrite("wall_k_bond", &AdhesionParams::wall_k_bond)
        .def_readwrite("wall_receptor_spacing", &AdhesionParams::wall_receptor_spacing)
        .def_readwrite("adhesion_matrix", &AdhesionParams::adhesion_matrix);
    py::class_<ShanChenParams>(m, "ShanChenParams")
        .def(py::init<>())
        .def_readwrite("enabled", &ShanChenParams::enabled)
        .def_readwrite("G", &ShanChenParams::G)
        .def_readwrite("rho0_sc", &ShanChenParams::rho0_sc)
        .def_readwrite("eos_type", &ShanChenParams::eos_type)
        .def_readwrite("cs_a", &ShanChenParams::cs_a)
        .def_readwrite("cs_b", &ShanChenParams::cs_b)
        .def_readwrite("cs_T", &ShanChenParams::cs_T)
        .def_readwrite("cs_R", &ShanChenParams::cs_R)
        .def_readwrite("wetting_angle", &ShanChenParams::wetting_angle)
        .def_readwrite("n_components", &ShanChenParams::n_components)
        .def_readwrite("G_12", &ShanChenParams::G_12)
        .def_readwrite("tau_2", &ShanChenParams::tau_2);

    py::class_<ScalarParams>(m, "ScalarParams")
        .def(py::init<>())
        .def_readwrite("enabled", &ScalarParams::enabled)
        .def_readwrite("n_species", &ScalarParams::n_species)
        .def_readwrite("diffusivity", &ScalarParams::diffusivity)
        .def_readwrite("inlet_concentration", &ScalarParams::inlet_concentration)
        .def_readwrite("periodic_y", &ScalarParams::periodic_y);

    // output configuration (io/output_config.h; python/bindings.cpp:720-801)
    py::class_<FluidOutputFields>(m, "FluidOutputFields")
        .def(py::init<>())
        .def_readwrite("density", &FluidOutputFields::density)
        .def_readwrite("velocity", &FluidOutputFields::velocity)
        .def_readwrite("pressure", &FluidOutputFields::pressure)
        .def_readwrite("vorticity", &FluidOutputFields::vorticity)
        .def_readwrite("strain_rate", &FluidOutputFields::strain_rate)
        .def_readwrite("ibm_force", &FluidOutputFields::ibm_force)
        .def_readwrite("node_type", &FluidOutputFields::node_type)
        .def_readwrite("velocity_mag", &FluidOutputFields::velocity_mag)
        .def_readwrite("component_density", &FluidOutputFields::component_density)
        .def_readwrite("concentration", &FluidOutputFields::concentration);
    py::class_<ParticleVTKFields>(m, "ParticleVTKFields")
        .def(py::init<>())
        .def_readwrite("velocity", &ParticleVTKFields::velocity)
        .def_readwrite("force", &ParticleVTKFields::force)
        .def_readwrite("particle_id", &ParticleVTKFields::particle_id)
        .def_readwrite("particle_type", &ParticleVTKFields::particle_type)
        .def_readwrite("particle_group", &ParticleVTKFields::particle_group)
        .def_readwrite("local_strain", &ParticleVTKFields::local_strain)
        .def_readwrite("local_curvature", &ParticleVTKFields::local_curvature);
    py::class_<ParticleOutputFields>(m, "ParticleOutputFields")
        .def(py::init<>())
        .def_readwrite("position", &ParticleOutputFields::position)
        .def_readwrite("velocity", &ParticleOutputFields::velocity)
        .def_readwrite("force", &ParticleOutputFields::force)
        .def_readwrite("diameter", &ParticleOutputFields::diameter)
        .def_readwrite("type", &ParticleOutputFields::type)
        .def_readwrite("group", &ParticleOutputFields::group)
        .def_readwrite("deformation", &ParticleOutputFields::deformation)
        .def_readwrite("area_volume", &ParticleOutputFields::area_volume)
        .def_readwrite("orientation", &ParticleOutputFields::orientation)
        .def_readwrite("angular_vel", &ParticleOutputFields::angular_vel)
        .def_readwrite("com_position", &ParticleOutputFields::com_position)
        .def_readwrite("com_velocity", &ParticleOutputFields::com_velocity);
    py::enum_<ParticleFilter::FilterType>(m, "FilterType")
        .value("ALL", ParticleFilter::ALL)
        .value("BY_ID", ParticleFilter::BY_ID)
        .value("BY_TYPE", ParticleFilter::BY_TYPE)
        .value("BY_GROUP", ParticleFilter::BY_GROUP)
        .value("BY_REGION", ParticleFilter::BY_REGION)
        .value("COMBINED", ParticleFilter::COMBINED);
    py::class_<ParticleFilter>(m, "ParticleFilter")
        .def(py::init<>())
        .def_readwrite("filter_type", &ParticleFilter::filter_type)
        .def_readwrite("selected_ids", &ParticleFilter::selected_ids)
        .def_readwrite("selected_types", &ParticleFilter::selected_types)
        .def_readwrite("selected_groups", &ParticleFilter::selected_groups)
        .def_readwrite("region_xmin", &ParticleFilter::region_xmin)
        .def_readwrite("region_xmax", &ParticleFilter::region_xmax)
        .def_readwrite("region_ymin", &ParticleFilter::region_ymin)
        .def_readwrite("region_ymax", &ParticleFilter::region_ymax)
        .def("passes", &ParticleFilter::passes, py::arg("id"), py::arg("type_id"), py::arg("group"),
             py::arg("cx"), py::arg("cy"));
    py::class_<FluidProbe>(m, "FluidProbe")
        .def(py::init<>())
        .def_readwrite("i", &FluidProbe::i)
        .def_readwrite("j", &FluidProbe::j)
        .def_readwrite("label", &FluidProbe::label);
    py::class_<ExtraCSVConfig>(m, "ExtraCSVConfig")
        .def(py::init<>())
        .def_readwrite("filename", &ExtraCSVConfig::filename)
        .def_readwrite("dump_every", &ExtraCSVConfig::dump_every)
        .def_readwrite("fields", &ExtraCSVConfig::fields)
        .def_readwrite("filter", &ExtraCSVConfig::filter)
        .def_readwrite("append", &ExtraCSVConfig::append);
    py::class_<OutputConfig>(m, "OutputConfig")
        .def(py::init<>())
        .def_readwrite("output_dir", &OutputConfig::output_dir)
        .def_readwrite("vtk_dump_every", &OutputConfig::vtk_dump_every)
        .def_readwrite("vtk_format", &OutputConfig::vtk_format)
        .def_readwrite("csv_dump_every", &OutputConfig::csv_dump_every)
        .def_readwrite("csv_format", &OutputConfig::csv_format)
        .def_readwrite("csv_append", &OutputConfig::csv_append)
        .def_readwrite("probe_dump_every", &OutputConfig::probe_dump_every)
        .def_readwrite("stats_dump_every", &OutputConfig::stats_dump_every)
        .def_readwrite("fluid_fields", &OutputConfig::fluid_fields)
        .def_readwrite("particle_vtk_fields", &OutputConfig::particle_vtk_fields)
        .def_readwrite("csv_fields", &OutputConfig::csv_fields)
        .def_readwrite("csv_filter", &OutputConfig::csv_filter)
        .def_readwrite("extra_csv", &OutputConfig::extra_csv)
        .def_readwrite("probes", &OutputConfig::probes);

    py::class_<SimulationParams>(m, "SimulationParams")
        .def(py::init<>())
        .def_readwrite("nx", &SimulationParams::nx)
        .def_readwrite("ny", &SimulationParams::ny)
        .def_readwrite("dt", &SimulationParams::dt)
        .def_readwrite("num_steps", &SimulationParams::num_steps)
        .def_readwrite("output_interval", &SimulationParams::output_interval)
        .def_readwrite("ibm_iterations", &SimulationParams::ibm_iterations)
        .def_readwrite("fluid", &SimulationParams::fluid)
        .def_readwrite("membrane", &SimulationParams::membrane)
        .def_readwrite("repulsion", &SimulationParams::repulsion)
        .def_readwrite("lubrication", &SimulationParams::lubrication)
        .def_readwrite("adhesion", &SimulationParams::adhesion)
        .def_readwrite("shan_chen", &SimulationParams::shan_chen)
        .def_readwrite("output_format", &SimulationParams::output_format)
        .def_readwrite("scalar", &SimulationParams::scalar)
        .def_readwrite("output_dir", &SimulationParams::output_dir)
        .def_readwrite("vtk_format", &SimulationParams::vtk_format)
        .def_readwrite("checkpoint_interval", &SimulationParams::checkpoint_interval)
        .def_readwrite("metrics_interval", &SimulationParams::metrics_interval)
        .def_readwrite("vtk_dump_every", &SimulationParams::vtk_dump_every)
        .def_readwrite("csv_dump_every", &SimulationParams::csv_dump_every)
        .def_readwrite("probe_dump_every", &SimulationParams::probe_dump_every)
        .def_readwrite("stats_dump_every", &SimulationParams::stats_dump_every)
        .def_readwrite("enable_stability_checks", &SimulationParams::enable_stability_checks)
        .def_readwrite("max_density_deviation", &SimulationParams::max_density_deviation)
        .def_readwrite("max_velocity", &SimulationParams::max_velocity)
        .def_readwrite("stability_check_interval", &SimulationParams::stability_check_interval)
        .def_readwrite("enable_profiling", &SimulationParams::enable_profiling)
        .def_readwrite("rng_seed", &SimulationParams::rng_seed)
        .def_readwrite("max_lattice_force", &SimulationParams::max_lattice_force)
        .def_readwrite("precision", &SimulationParams::precision)
        .def_readwrite("use_graph", &SimulationParams::use_graph)
        .def_readwrite("device", &SimulationParams::device)
        .def("kinematicViscosity", &SimulationParams::kinematicViscosity);

    py::class_<LatticeField>(m, "LatticeField")
        .def("getNx", &LatticeField::getNx)
        .def("getNy", &LatticeField::getNy)
        .def("size", &LatticeField::size)
        .def("density", [](py::object self) { return view2d<LatticeField>(self, LatticeField::VIEW_RHO); })
        .def("velocity_x", [](py::object self) { return view2d<LatticeField>(self, LatticeField::VIEW_UX); })
        .def("velocity_y", [](py::object self) { return view2d<LatticeField>(self, LatticeField::VIEW_UY); })
        .def("force_x", [](py::object self) { return view2d<LatticeField>(self, LatticeField::VIEW_FX); })
        .def("force_y", [](py::object self) { return view2d<LatticeField>(self, LatticeField::VIEW_FY); })
        .def("distributions", [](py::object self) {
            LatticeField& f = self.cast<LatticeField&>();
            Real* data = nullptr;
            py::capsule owner = lease<LatticeField>(self, LatticeField::VIEW_F, &data);
            const py::ssize_t N = f.size(), nx = f.getNx(), s = sizeof(Real);
            return py::array(py::dtype::of<Real>(), {static_cast<py::ssize_t>(9), static_cast<py::ssize_t>(f.getNy()), nx},
                             {N * s, nx * s, s}, data, owner);
        })
        .def("cell_types", [](LatticeField& f) {
            py::array_t<int> a({f.getNy(), f.getNx()});
            std::memcpy(a.mutable_data(), f.cellTypeData(), sizeof(int) * f.size());
            return a;
        })
        .def("getRho", &LatticeField::getRho, py::arg("x"), py::arg("y"))
        .def("getUx", &LatticeField::getUx, py::arg("x"), py::arg("y"))
        .def("getUy", &LatticeField::getUy, py::arg("x"), py::arg("y"))
        .def("getCellType", &LatticeField::getCellType, py::arg("x"), py::arg("y"))
        .def("setCellType", &LatticeField::setCellType, py::arg("x"), py::arg("y"), py::arg("ct"))
        .def("setRho", &LatticeField::setRho, py::arg("x"), py::arg("y"), py::arg("rho"))
        .def("setUx", &LatticeField::setUx, py::arg("x"), py::arg("y"), py::arg("ux"))
        .def("setUy", &LatticeField::setUy, py::arg("x"), py::arg("y"), py::arg("uy"))
        .def("setEquilibrium", &LatticeField::setEquilibrium, py::arg("rho0"), py::arg("ux0"), py::arg("uy0"))
        .def("initializeEquilibriumAt", &LatticeField::initializeEquilibriumAt, py::arg("x"), py::arg("y"))
        .def("initializeAllEquilibrium", &LatticeField::initializeAllEquilibrium)
        .def("computeMacroscopic", &LatticeField::computeMacroscopic)
        .def("f", &LatticeField::f, py::arg("x"), py::arg("y"), py::arg("q"))
        .def("set_f", &LatticeField::set_f, py::arg("x"), py::arg("y"), py::arg("q"), py::arg("value"));

    py::class_<Capsule>(m, "Capsule")
        .def("nodePosition", &Capsule::nodePosition, py::arg("k"))
        .def("nodeVelocity", &Capsule::nodeVelocity, py::arg("k"))
        .def("nodeForce", &Capsule::nodeForce, py::arg("k"))
        .def("setNodeVelocity", &Capsule::setNodeVelocity, py::arg("k"), py::arg("v"))
        .def("centroid", &Capsule::centroid)
        .def("area", &Capsule::area)
        .def("perimeter", &Capsule::perimeter)
        .def("deformationIndex", &Capsule::deformationIndex)
        .def("inclinationAngle", &Capsule::inclinationAngle)
        .def("effectiveRadius", &Capsule::effectiveRadius)
        .def("numNodes", &Capsule::numNodes)
        .def("getId", &Capsule::getId)
        .def("getType", &Capsule::getType)
        .def("getTypeId", &Capsule::getType)
        .def("setPeriodicX", &Capsule::setPeriodicX, py::arg("nx"))
        .def("getPeriodicX", &Capsule::getPeriodicX)
        .def("node_positions_array", [](const Capsule& c) {
            const int n = c.numNodes();
            py::array_t<Real> a({n, 2});
            auto b = a.mutable_unchecked<2>();
            for (int i = 0; i < n; ++i) { Vec2d p = c.nodePosition(i); b(i, 0) = p.x; b(i, 1) = p.y; }
            return a;
        });

    py::class_<CapsuleSystem>(m, "CapsuleSystem")
        .def("addCapsule", &CapsuleSystem::addCapsule, py::arg("center"), py::arg("radius"),
             py::arg("num_nodes") = 0, py::arg("params"), py::arg("type") = 0)
        .def("numCapsules", &CapsuleSystem::numCapsules)
        .def("totalNodes", &CapsuleSystem::totalNodes)
        .def("__getitem__", [](CapsuleSystem& cs, int i) -> Capsule& {
            if (i < 0 || i >= cs.numCapsules()) throw py::index_error("Capsule index out of range");
            return cs[i];
        }, py::return_value_policy::reference_internal)
        .def("__len__", &CapsuleSystem::numCapsules);

    py::class_<AdvectionDiffusion>(m, "AdvectionDiffusion")
        .def("initialize", &AdvectionDiffusion::initialize, py::arg("initial_concentration") = 0.0)
        .def("setRegion", &AdvectionDiffusion::setRegion, py::arg("x0"), py::arg("y0"), py::arg("x1"),
             py::arg("y1"), py::arg("concentration"), py::arg("species") = 0)
        .def("setPoint", &AdvectionDiffusion::setPoint, py::arg("x"), py::arg("y"),
             py::arg("concentration"), py::arg("species") = 0)
        .def("getConcentration", &AdvectionDiffusion::getConcentration, py::arg("x"), py::arg("y"),
             py::arg("species") = 0)
        .def("concentration", [](py::object self, int species) {
            AdvectionDiffusion& a = self.cast<AdvectionDiffusion&>();
            if (species < 0 || species >= a.getNumSpecies()) throw py::index_error("no such species");
            return view2d<AdvectionDiffusion>(self, species);
        }, py::arg("species") = 0)
        .def("getNumSpecies", &AdvectionDiffusion::getNumSpecies)
        .def("setField", [](AdvectionDiffusion& a, py::array_t<Real, py::array::c_style | py::array::forcecast> C,
                            int species) {
            if (C.size() != static_cast<py::ssize_t>(a.getNx()) * a.getNy())
                throw py::value_error("concentration array must have ny*nx entries");
            a.setField(C.data(), species);
        }, py::arg("C"), py::arg("species") = 0);

    // geometry: obstacles and the channel builder (python/bindings.cpp:602-652)
    py::class_<Obstacle, std::shared_ptr<Obstacle>>(m, "Obstacle")
        .def("contains", &Obstacle::contains, py::arg("x"), py::arg("y"))
        .def("signedDistance", &Obstacle::signedDistance, py::arg("x"), py::arg("y"))
        .def("nearestPoint", &Obstacle::nearestPoint, py::arg("x"), py::arg("y"))
        .def("normalAt", &Obstacle::normalAt, py::arg("x"), py::arg("y"));
    py::class_<CircleObstacle, Obstacle, std::shared_ptr<CircleObstacle>>(m, "CircleObstacle")
        .def(py::init<Real, Real, Real>(), py::arg("cx"), py::arg("cy"), py::arg("radius"))
        .def("getCx", &CircleObstacle::getCx)
        .def("getCy", &CircleObstacle::getCy)
        .def("getRadius", &CircleObstacle::getRadius);
    py::class_<RectObstacle, Obstacle, std::shared_ptr<RectObstacle>>(m, "RectObstacle")
        .def(py::init<Real, Real, Real, Real>(), py::arg("x0"), py::arg("y0"), py::arg("x1"), py::arg("y1"))
        .def("getX0", &RectObstacle::getX0)
        .def("getY0", &RectObstacle::getY0)
        .def("getX1", &RectObstacle::getX1)
        .def("getY1", &RectObstacle::getY1);
    py::class_<PolygonObstacle, Obstacle, std::shared_ptr<PolygonObstacle>>(m, "PolygonObstacle")
        .def(py::init<const std::vector<Vec2d>&>(), py::arg("vertices"));

    py::class_<ChannelBuilder>(m, "ChannelBuilder")
        .def(py::init<int, int>(), py::arg("nx"), py::arg("ny"))
        .def("addWalls", &ChannelBuilder::addWalls, py::return_value_policy::reference_internal)
        .def("addCirclePillar", &ChannelBuilder::addCirclePillar, py::arg("cx"), py::arg("cy"),
             py::arg("radius"), py::return_value_policy::reference_internal)
        .def("addRectPillar", &ChannelBuilder::addRectPillar, py::arg("x0"), py::arg("y0"), py::arg("x1"),
             py::arg("y1"), py::return_value_policy::reference_internal)
        .def("addPolygonPillar", &ChannelBuilder::addPolygonPillar, py::arg("vertices"),
             py::return_value_policy::reference_internal)
        .def("setBoundaryType", &ChannelBuilder::setBoundaryType, py::arg("bt"),
             py::return_value_policy::reference_internal)
        .def("getBoundaryType", &ChannelBuilder::getBoundaryType)
        .def("hasTopWall", &ChannelBuilder::hasTopWall)
        .def("hasBottomWall", &ChannelBuilder::hasBottomWall)
        .def("getNx", &ChannelBuilder::getNx)
        .def("getNy", &ChannelBuilder::getNy)
        .def("getObstacles", &ChannelBuilder::getObstacles)
        .def("wallBottom", &ChannelBuilder::wallBottom)
        .def("wallTop", &ChannelBuilder::wallTop);

    py::class_<LBMSolverView>(m, "LBMSolver")
        .def("field", &LBMSolverView::field, py::return_value_policy::reference_internal)
        .def("getNx", &LBMSolverView::getNx)
        .def("getNy", &LBMSolverView::getNy)
        .def("getStep", &LBMSolverView::getStep)
        .def("isPeriodicX", &LBMSolverView::isPeriodicX)
        .def("isPeriodicY", &LBMSolverView::isPeriodicY);

    py::class_<Simulation>(m, "Simulation")
        .def(py::init<const SimulationParams&>(), py::arg("params"))
        .def("setChannelBuilder", &Simulation::setChannelBuilder, py::arg("builder"))
        .def("addCapsule", &Simulation::addCapsule, py::arg("center"), py::arg("radius"),
             py::arg("num_nodes") = 0, py::arg("mparams"), py::arg("type") = 0)
        .def("addCapsuleRandom", &Simulation::addCapsuleRandom, py::arg("count"), py::arg("x0"),
             py::arg("y0"), py::arg("x1"), py::arg("y1"), py::arg("radius_min"), py::arg("radius_max"),
             py::arg("num_nodes") = 0, py::arg("mparams"), py::arg("type") = 0,
             py::arg("seed") = 12345, py::arg("min_gap") = 1.0, py::arg("max_attempts") = 0)
        .def("initialize", &Simulation::initialize)
        .def("step", &Simulation::step)
        .def("finalize", &Simulation::finalize, py::call_guard<py::gil_scoped_release>())
        .def("setOutputConfig", &Simulation::setOutputConfig, py::arg("config"))
        .def("setVTKOutput", &Simulation::setVTKOutput, py::arg("dump_every"), py::arg("format"),
             py::arg("fluid_fields"), py::arg("particle_fields"))
        .def("setCSVOutput", &Simulation::setCSVOutput, py::arg("dump_every"), py::arg("format"),
             py::arg("fields"), py::arg("filter"), py::arg("append") = true)
        .def("addExtraCSVOutput", &Simulation::addExtraCSVOutput, py::arg("filename"), py::arg("dump_every"),
             py::arg("fields"), py::arg("filter"))
        .def("addFluidProbe", &Simulation::addFluidProbe, py::arg("i"), py::arg("j"), py::arg("label"))
        .def("run", &Simulation::run, py::arg("num_steps"), py::call_guard<py::gil_scoped_release>())
        .def("currentStep", &Simulation::currentStep)
        .def("getMaxSpeed", &Simulation::getMaxSpeed)
        .def("lastCappedNodes", &Simulation::lastCappedNodes)
        .def("checkStability", &Simulation::checkStability)
        .def("fieldStats", &Simulation::fieldStats)
        .def("globalStats", [](Simulation& s) {
            py::dict d;
            for (auto& kv : s.globalStats()) d[py::str(kv.first)] = kv.second;
            return d;
        })
        .def("saveCheckpoint", &Simulation::saveCheckpoint, py::arg("filename"))
        .def("loadCheckpoint", &Simulation::loadCheckpoint, py::arg("filename"))
        .def("setCurrentStep", &Simulation::setCurrentStep, py::arg("step"))
        .def("lbmSolver", [](Simulation& s) { return LBMSolverView{&s}; }, py::keep_alive<0, 1>())
        .def("field", [](Simulation& s) -> LatticeField& { return s.field(); },
             py::return_value_policy::reference_internal)
        .def("capsules", [](Simulation& s) -> CapsuleSystem& { return s.capsules(); },
             py::return_value_policy::reference_internal)
        .def("params", [](const Simulation& s) -> const SimulationParams& { return s.params(); },
             py::return_value_policy::reference_internal)
        .def("advectionDiffusion", [](Simulation& s) { return s.advectionDiffusion(); },
             py::return_value_policy::reference_internal)
        .def("setFluidRegion", &Simulation::setFluidRegion, py::arg("x0"), py::arg("y0"), py::arg("x1"),
             py::arg("y1"), py::arg("rho"), py::arg("ux"), py::arg("uy"))
        .def("setScalarRegion", &Simulation::setScalarRegion, py::arg("x0"), py::arg("y0"),
             py::arg("x1"), py::arg("y1"), py::arg("concentration"), py::arg("species") = 0)
        .def("setStepCallback", [](Simulation& sim, py::function cb) {
            sim.setStepCallback([cb](Simulation& s, int step) {
                py::gil_scoped_acquire acquire;
                cb(py::cast(s, py::return_value_policy::reference), step);
            });
        }, py::arg("callback"))
        .def("setTiming", &Simulation::setTiming, py::arg("on"))
        .def("synchronize", &Simulation::synchronize, py::call_guard<py::gil_scoped_release>())
        .def("graphCaptures", &Simulation::graphCaptures)
        .def("autotune", &Simulation::autotune, py::arg("reps") = 10)
        .def("setLbmPairs", &Simulation::setLbmPairs, py::arg("on"))
        .def("setCapsuleLaunch", &Simulation::setCapsuleLaunch, py::arg("spread_grid"), py::arg("interp_lanes"))
        .def("setLbmLaunch", &Simulation::setLbmLaunch, py::arg("bx"), py::arg("by"), py::arg("xcd_banded") = true,
             py::arg("persist") = 0)
        .def("kernelTimes", &Simulation::kernelTimes)
        .def("setMdfControl", &Simulation::setMdfControl, py::arg("tolerance"),
             py::arg("max_corrections") = 8, py::arg("probe") = false)
        .def("setMdfTail", &Simulation::setMdfTail, py::arg("first"))
        .def("mdfStats", [](Simulation& s) {
            auto v = s.mdfStats();
            py::dict d;
            d["mean_corrections"] = v[0];
            d["mean_residual"] = v[1];
            d["max_residual"] = v[2];
            d["steps"] = v[3];
            return d;
        })
        .def("collectMdfSamples", &Simulation::collectMdfSamples, py::arg("capacity"))
        .def("mdfSampleCount", &Simulation::mdfSampleCount)
        .def("mdfSamples", [](Simulation& s) {
            auto v = s.mdfSamples();
            const py::ssize_t w = SF_MLP_IN + SF_MLP_OUT;
            py::array_t<float> a({static_cast<py::ssize_t>(v.size()) / w, w});
            std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(float));
            return a;
        })
        .def("setMdfModel", &Simulation::setMdfModel, py::arg("model"), py::arg("warm") = true,
             py::keep_alive<1, 2>());

    py::class_<Ensemble>(m, "Ensemble")
        .def(py::init<const std::vector<Simulation*>&>(), py::arg("members"), py::keep_alive<1, 2>())
        .def("run", &Ensemble::run, py::arg("num_steps"), py::call_guard<py::gil_scoped_release>())
        .def("synchronize", &Ensemble::synchronize)
        .def("autotune", &Ensemble::autotune, py::arg("reps") = 10)
        .def("setLbmLaunch", &Ensemble::setLbmLaunch, py::arg("bx"), py::arg("by"), py::arg("xcd_banded") = true)
        .def("setLbmFlat", &Ensemble::setLbmFlat, py::arg("on"))
        .def("setGraphMode", &Ensemble::setGraphMode, py::arg("on"))
        .def("info", &Ensemble::info)
        .def("__len__", &Ensemble::size);

    py::class_<MdfModel>(m, "MdfModel")
        .def(py::init<int, uint64_t>(), py::arg("device") = 0, py::arg("seed") = 1)
        .def("train", &MdfModel::train, py::arg("sim"), py::arg("epochs") = 10, py::arg("batch") = 4096,
             py::arg("lr") = 1e-3, py::arg("seed") = 1, py::call_guard<py::gil_scoped_release>())
        .def("params", [](const MdfModel& m) { auto v = m.params(); return py::array_t<float>(v.size(), v.data()); })
        .def("normalization", [](const MdfModel& m) { auto v = m.normalization(); return py::array_t<float>(v.size(), v.data()); })
        .def("setParams", [](MdfModel& m, py::array_t<float, py::array::c_style | py::array::forcecast> p) {
            m.setParams(std::vector<float>(p.data(), p.data() + p.size()));
        })
        .def("setNormalization", [](MdfModel& m, py::array_t<float, py::array::c_style | py::array::forcecast> n) {
            m.setNormalization(std::vector<float>(n.data(), n.data() + n.size()));
        })
        .def("predict", [](const MdfModel& m, py::array_t<float, py::array::c_style | py::array::forcecast> x) {
            auto y = m.predict(std::vector<float>(x.data(), x.data() + x.size()));
            py::array_t<float> a({static_cast<py::ssize_t>(y.size() / SF_MLP_OUT), static_cast<py::ssize_t>(SF_MLP_OUT)});
            std::memcpy(a.mutable_data(), y.data(), y.size() * sizeof(float));
            return a;
        })
        .def("lossGrad", [](const MdfModel& m, py::array_t<float, py::array::c_style | py::array::forcecast> x,
                            py::array_t<float, py::array::c_style | py::array::forcecast> t) {
            auto r = m.lossGrad(std::vector<float>(x.data(), x.data() + x.size()),
                                std::vector<float>(t.data(), t.data() + t.size()));
            return py::make_tuple(r.first, py::array_t<float>(r.second.size(), r.second.data()));
        });

    m.def("device_count", []() {
        int n = 0;
        if (sf_device_count(&n) != SF_OK) return 0;
        return n;
    });
}
