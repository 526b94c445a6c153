// svm.hpp -- C++ adapter over the C ABI (svm.h) that mirrors the reference's
// class API, so a driver written against the reference's classes keeps its
// shape:
//
//   simpleModel                    src/model/simpleModel.h:33-145
//   sphericalVertexModel           src/model/sphericalVertexModel.h:10-85
//   sphericalVoronoi               src/model/sphericalVoronoi.h (cell points as
//                                  the particles, retriangulated after every
//                                  move), with forces of the vertex-model
//                                  energy of its Voronoi cells
//   brownianDynamics               src/updaters/brownianDynamics.h:8-28
//   sphericalSelfPropelledParticle src/updaters/sphericalSelfPropelledParticle.h:13-22,
//                                  src/updaters/vectorialVicsek.h:31-42
//   energyMinimizerFIRE            src/updaters/energyMinimizerFIRE.h:21-96
//   Simulation                     src/simulation/simulation.h:18-77
//   neighborList                   src/utilities/neighborList.h:8-76
//
// Differences a caller sees: the model is built from a seed (the reference
// takes a noiseSource) and always runs on the GPU (there is no CPU path in
// this library; the CPU restatement is test infrastructure); state lives on
// the device and return* methods copy it out.  Errors throw std::runtime_error
// with the library's message, as the reference's HANDLE_ERROR throws.
// Header-only; link with libsvm.so.
#pragma once
#include <array>
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "svm.h"

namespace svm {

using dVec = std::array<double, 3>;

inline void check(int rc, svm_context* ctx, const char* what) {
    if (rc != SVM_OK) {
        std::string msg = std::string(what) + " failed (status " + std::to_string(rc) + ")";
        if (ctx) msg += ": " + std::string(svm_last_error(ctx));
        throw std::runtime_error(msg);
    }
}

// What the updaters and the Simulation see of a model (simpleModel,
// src/model/simpleModel.h:33-145).
// Metric neighbour lists of a set of points (neighborList,
// src/utilities/neighborList.h:8-76): neighborList(range, sphereRadius), then
// computeNeighborLists(points) fills neighborsPerParticle, particleIndices,
// neighborVectors and neighborDistances, entry n of particle p at
// p * Nmax + n (the reference's neighborIndexer(n, p)); Nmax grows and the
// lists are recomputed when a particle has more neighbours, as in the
// reference (:150-153).  Computed on the GPU (svm_neighbor_list); each
// particle's neighbours in increasing index order.
class neighborList {
   public:
    neighborList(double range, double sphereRadius) : maxRange(range), radius(sphereRadius) {}
    void computeNeighborLists(const std::vector<dVec>& points) {
        const int n = (int)points.size();
        std::vector<double> p(3 * (size_t)n);
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) p[3 * (size_t)i + k] = points[i][k];
        neighborsPerParticle.assign(n, 0);
        while (true) {
            particleIndices.assign((size_t)n * Nmax, 0);
            std::vector<double> vec(3 * (size_t)n * Nmax), dist((size_t)n * Nmax);
            int32_t needed = 0;
            const int rc = svm_neighbor_list(n, p.data(), maxRange, radius, Nmax, neighborsPerParticle.data(),
                                             particleIndices.data(), vec.data(), dist.data(), &needed);
            if (rc == SVM_ERR_CAPACITY && needed > Nmax) {
                Nmax = needed;
                continue;
            }
            check(rc, nullptr, "svm_neighbor_list");
            neighborVectors.assign((size_t)n * Nmax, dVec{0.0, 0.0, 0.0});
            for (size_t e = 0; e < (size_t)n * Nmax; ++e)
                neighborVectors[e] = dVec{vec[3 * e], vec[3 * e + 1], vec[3 * e + 2]};
            neighborDistances = dist;
            return;
        }
    }
    double maxRange, radius;
    int Nmax = 4;   // neighborList::setBasics (src/utilities/neighborList.cpp:18)
    std::vector<int32_t> neighborsPerParticle;
    std::vector<int32_t> particleIndices;
    std::vector<dVec> neighborVectors;
    std::vector<double> neighborDistances;
};

class simpleModel {
   public:
    virtual ~simpleModel() = default;
    virtual svm_context* context() = 0;
    // true for sphericalVoronoi: the cell points are the degrees of freedom
    virtual bool voronoi() const { return false; }
    virtual int getNumberOfParticles() const = 0;
    virtual void computeGeometry() = 0;
    virtual void computeForces(bool zeroOutForces = false) = 0;
    virtual double computeEnergy() = 0;
    virtual double computeKineticEnergy() = 0;
};

// One sphere.  sphericalVertexModel(n, noise, area, perimeter, useGPU, neverGPU)
// (src/model/sphericalVertexModel.cpp:5-82): seeded uniform cell points on a
// sphere of mean cell area 1, their Delaunay topology (built on the GPU),
// seeded unit tangent directors.
class sphericalVertexModel : public simpleModel {
   public:
    sphericalVertexModel(int nCells, uint64_t seed, double area = 1.0, double perimeter = 3.8, double kr = 1.0,
                         double t1Threshold = 0.1)
        : nCells_(nCells) {
        svm_system_desc d{};
        d.n_cells = nCells;
        d.ring_capacity = 32;
        d.a0 = area;
        d.p0 = perimeter;
        d.kr = kr;
        d.t1_threshold = t1Threshold;
        d.seed = seed;
        check(svm_create(&d, 1, &ctx_), nullptr, "svm_create");
        std::vector<double> pts(3 * (size_t)nCells);
        check(svm_generate_cell_points(nCells, seed, pts.data()), ctx_, "svm_generate_cell_points");
        check(svm_build_from_cell_points(ctx_, 0, pts.data()), ctx_, "svm_build_from_cell_points");
        int32_t nv = 0;
        check(svm_get_num_vertices(ctx_, 0, &nv), ctx_, "svm_get_num_vertices");
        N_ = nv;
        std::vector<double> pos(3 * (size_t)N_), dir(3 * (size_t)N_);
        check(svm_get_positions(ctx_, 0, pos.data()), ctx_, "svm_get_positions");
        check(svm_generate_directors(N_, pos.data(), seed, dir.data()), ctx_, "svm_generate_directors");
        check(svm_set_directors(ctx_, 0, dir.data()), ctx_, "svm_set_directors");
    }
    ~sphericalVertexModel() {
        if (ctx_) svm_destroy(ctx_);
    }
    sphericalVertexModel(const sphericalVertexModel&) = delete;
    sphericalVertexModel& operator=(const sphericalVertexModel&) = delete;

    // simpleModel::getNumberOfParticles: the vertices are the particles.
    int getNumberOfParticles() const override { return N_; }
    int getNumberOfCells() const { return nCells_; }
    void computeGeometry() override { check(svm_compute_geometry(ctx_), ctx_, "svm_compute_geometry"); }
    void computeForces(bool zeroOutForces = false) override {
        (void)zeroOutForces;   // the self force always overwrites
        check(svm_compute_forces(ctx_, forcePrecision_), ctx_, "svm_compute_forces");
    }
    void moveParticles(const std::vector<dVec>& displacements, double scale = 1.) {
        check(svm_move_particles(ctx_, 0, &displacements[0][0], scale), ctx_, "svm_move_particles");
    }
    void enforceTopology() { check(svm_enforce_topology(ctx_), ctx_, "svm_enforce_topology"); }
    // sphericalVertexModel::computeEnergy: sum Kr (A - A0)^2 + (P - P0)^2 of
    // the last computed geometry.
    double computeEnergy() override {
        double pe = 0.0, ke = 0.0;
        check(svm_get_energies(ctx_, &pe, &ke), ctx_, "svm_get_energies");
        return pe;
    }
    double computeKineticEnergy() override {
        double pe = 0.0, ke = 0.0;
        check(svm_get_energies(ctx_, &pe, &ke), ctx_, "svm_get_energies");
        return ke;
    }
    void setPreferredParameters(double a0, double p0) {
        check(svm_set_preferred_parameters(ctx_, 0, a0, p0), ctx_, "svm_set_preferred_parameters");
    }
    void setScalarModelParameter(double kr) {
        check(svm_set_scalar_model_parameter(ctx_, 0, kr), ctx_, "svm_set_scalar_model_parameter");
    }
    void setT1Threshold(double t) { check(svm_set_t1_threshold(ctx_, 0, t), ctx_, "svm_set_t1_threshold"); }
    // the SCALARFLOAT switch as a run-time choice: 0 FP64, 1 FP32 forces
    void setForcePrecision(int p) {
        forcePrecision_ = p;
        check(svm_set_force_precision(ctx_, p), ctx_, "svm_set_force_precision");
    }
    void setParticlePositions(const std::vector<dVec>& p) {
        check(svm_set_positions(ctx_, 0, &p[0][0]), ctx_, "svm_set_positions");
    }
    std::vector<dVec> returnPositions() { return get3(svm_get_positions, N_, "svm_get_positions"); }
    std::vector<dVec> returnForces() { return get3(svm_get_forces, N_, "svm_get_forces"); }
    std::vector<dVec> returnVelocities() { return get3(svm_get_velocities, N_, "svm_get_velocities"); }
    std::vector<dVec> returnDirectors() { return get3(svm_get_directors, N_, "svm_get_directors"); }
    std::vector<dVec> returnCellPositions() { return get3(svm_get_cell_positions, nCells_, "svm_get_cell_positions"); }
    std::vector<std::array<double, 2>> returnAreaPerimeter() {
        std::vector<std::array<double, 2>> out((size_t)nCells_);
        check(svm_get_area_perimeter(ctx_, 0, &out[0][0]), ctx_, "svm_get_area_perimeter");
        return out;
    }
    svm_context* context() override { return ctx_; }

   private:
    template <class F>
    std::vector<dVec> get3(F f, int n, const char* what) {
        std::vector<dVec> out((size_t)n);
        check(f(ctx_, 0, &out[0][0]), ctx_, what);
        return out;
    }
    svm_context* ctx_ = nullptr;
    int nCells_ = 0, N_ = 0;
    int forcePrecision_ = 0;
};

// One sphere of the Voronoi model (sphericalVoronoi, src/model/sphericalVoronoi.h;
// sphericalModel::setParticlePositionsRandomly, src/model/sphericalModel.cpp:37-65):
// seeded uniform cell points, the particles, with their Delaunay topology
// built on the GPU and rebuilt after every move.  The reference model has no
// forces; here they are those of E = sum Kr (A - A0)^2 + (P - P0)^2 over the
// Voronoi cells (circumcentre vertices), carried to the cell points.
class sphericalVoronoi : public simpleModel {
   public:
    sphericalVoronoi(int nCells, uint64_t seed, double area = 1.0, double perimeter = 3.8, double kr = 1.0)
        : N_(nCells) {
        svm_system_desc d{};
        d.n_cells = nCells;
        d.ring_capacity = 32;
        d.a0 = area;
        d.p0 = perimeter;
        d.kr = kr;
        d.t1_threshold = 0.1;
        d.seed = seed;
        check(svm_create(&d, 1, &ctx_), nullptr, "svm_create");
        std::vector<double> pts(3 * (size_t)nCells);
        check(svm_generate_cell_points(nCells, seed, pts.data()), ctx_, "svm_generate_cell_points");
        check(svm_build_voronoi_from_cell_points(ctx_, 0, pts.data()), ctx_, "svm_build_voronoi_from_cell_points");
    }
    ~sphericalVoronoi() {
        if (ctx_) svm_destroy(ctx_);
    }
    sphericalVoronoi(const sphericalVoronoi&) = delete;
    sphericalVoronoi& operator=(const sphericalVoronoi&) = delete;

    bool voronoi() const override { return true; }
    int getNumberOfParticles() const override { return N_; }
    void computeGeometry() override { check(svm_compute_geometry(ctx_), ctx_, "svm_compute_geometry"); }
    void computeForces(bool zeroOutForces = false) override {
        (void)zeroOutForces;
        check(svm_voronoi_forces(ctx_), ctx_, "svm_voronoi_forces");
    }
    double computeEnergy() override {
        double pe = 0.0, ke = 0.0;
        check(svm_get_energies(ctx_, &pe, &ke), ctx_, "svm_get_energies");
        return pe;
    }
    double computeKineticEnergy() override { return 0.0; }
    void setPreferredParameters(double a0, double p0) {
        check(svm_set_preferred_parameters(ctx_, 0, a0, p0), ctx_, "svm_set_preferred_parameters");
    }
    // the cell points, and the forces on them of the last computeForces
    std::vector<dVec> returnPositions() { return get3(svm_get_generators, "svm_get_generators"); }
    std::vector<dVec> returnForces() { return get3(svm_get_voronoi_forces, "svm_get_voronoi_forces"); }
    svm_context* context() override { return ctx_; }

   private:
    template <class F>
    std::vector<dVec> get3(F f, const char* what) {
        std::vector<dVec> out((size_t)N_);
        check(f(ctx_, 0, &out[0][0]), ctx_, what);
        return out;
    }
    svm_context* ctx_ = nullptr;
    int N_ = 0;
};

// updater (src/updaters/baseUpdater.h:19-102)
class updater {
   public:
    virtual ~updater() = default;
    virtual void setModel(std::shared_ptr<simpleModel> m) { model = std::move(m); }
    virtual void setDeltaT(double dt) { deltaT = dt; }
    virtual void setMaximumIterations(int maxIt) { maxIterations = maxIt; }
    virtual int getCurrentIterations() { return iterations; }
    virtual void performUpdate() = 0;

   protected:
    // the updaters below other than brownianDynamics move vertex models only
    svm_context* vertexModelContext(const char* who) {
        if (model->voronoi()) throw std::runtime_error(std::string(who) + " needs a sphericalVertexModel");
        return model->context();
    }
    std::shared_ptr<simpleModel> model;
    double deltaT = 0.01;
    int maxIterations = 1000;
    int iterations = 0;
};
typedef std::shared_ptr<updater> UpdaterPtr;

// brownianDynamics: d = sqrt(2 dt mu T) xi + dt mu F (brownianDynamics.cpp:29-48),
// on the vertices of a vertex model or the cell points of a Voronoi model
// (which are then retriangulated).
class brownianDynamics : public updater {
   public:
    explicit brownianDynamics(bool reproducible = true) { (void)reproducible; }
    void setT(double t) { temperature = t; }
    void setMu(double m) { mu = m; }
    void performUpdate() override {
        svm_context* c = model->context();
        check(svm_set_brownian(c, 0, deltaT, mu, temperature), c, "svm_set_brownian");
        if (model->voronoi()) check(svm_voronoi_brownian_steps(c, 1), c, "svm_voronoi_brownian_steps");
        else check(svm_brownian_steps(c, 1), c, "svm_brownian_steps");
    }

   private:
    double temperature = 0.0, mu = 1.0;
};

// sphericalSelfPropelledParticle (sphericalSelfPropelledParticle.cpp:16-47)
class sphericalSelfPropelledParticle : public updater {
   public:
    void setMu(double m) { mu = m; }
    void setV0(double v) { v0 = v; }
    void setTau(double t) { tau = t; }
    void performUpdate() override {
        svm_context* c = vertexModelContext("sphericalSelfPropelledParticle");
        check(svm_set_active(c, 0, deltaT, mu, v0, tau), c, "svm_set_active");
        check(svm_active_steps(c, 1), c, "svm_active_steps");
    }

   private:
    double mu = 1.0, v0 = 0.1, tau = 1.0;
};

// sphericalVectorialVicsek (src/updaters/sphericalVectorialVicsek.h:8-27,
// setters of vectorialVicsek.h:31-34): alignment with the vertex neighbours.
class sphericalVectorialVicsek : public updater {
   public:
    void setMu(double m) { mu = m; }
    void setV0(double v) { v0 = v; }
    void setTau(double t) { tau = t; }
    void setEta(double e) { eta = e; }
    void performUpdate() override {
        svm_context* c = vertexModelContext("sphericalVectorialVicsek");
        check(svm_set_vicsek(c, 0, deltaT, mu, v0, tau, eta), c, "svm_set_vicsek");
        check(svm_vicsek_steps(c, 1), c, "svm_vicsek_steps");
    }

   private:
    double mu = 1.0, v0 = 0.01, tau = 1.0, eta = 1.0;
};

// energyMinimizerFIRE (energyMinimizerFIRE.cpp:58-273)
class energyMinimizerFIRE : public updater {
   public:
    explicit energyMinimizerFIRE(std::shared_ptr<sphericalVertexModel> m) { setModel(std::move(m)); }
    void setFIREParameters(double dt, double alphaStart, double deltaTMax, double deltaTInc, double deltaTDec,
                           double alphaDec, int nMin, double forceCutoff, double alphaMin = 0.75) {
        p_.dt = dt;
        p_.alpha_start = alphaStart;
        p_.dt_max = deltaTMax;
        p_.dt_inc = deltaTInc;
        p_.dt_dec = deltaTDec;
        p_.alpha_dec = alphaDec;
        p_.n_min = nMin;
        p_.force_cutoff = forceCutoff;
        p_.alpha_min = alphaMin;
        deltaT = dt;
        dirty_ = true;
    }
    // stop test: 0 = the reference's sqrt(sum F^2)/N, 1 = rms force
    void setStopRms(bool rms) {
        p_.stop_rms = rms ? 1 : 0;
        dirty_ = true;
    }
    void setMaximumIterations(int maxIt) override {
        maxIterations = maxIt;
        p_.max_iterations = maxIt;
        if (!dirty_) check(svm_fire_set_max_iterations(model->context(), 0, maxIt), model->context(),
                           "svm_fire_set_max_iterations");
    }
    void minimize() {
        svm_context* c = vertexModelContext("energyMinimizerFIRE");
        if (dirty_) {
            p_.max_iterations = maxIterations;
            check(svm_fire_set_parameters(c, 0, &p_), c, "svm_fire_set_parameters");
            dirty_ = false;
        }
        check(svm_fire_minimize(c, 100), c, "svm_fire_minimize");
        svm_fire_state s;
        check(svm_fire_get_state(c, 0, &s), c, "svm_fire_get_state");
        iterations = s.iterations;
        forceMax = s.force_max;
    }
    void performUpdate() override { minimize(); }
    double getMaxForce() const { return forceMax; }

   private:
    svm_fire_params p_{0.01, 0.99, 0.1, 1.05, 0.95, 0.9, 5, 1000, 1e-7, 0.75, 0, 0};
    bool dirty_ = true;
    double forceMax = 0.0;
};

// energyMinimizerAdam (src/updaters/energyMinimizerAdam.h:8-62).  Parameters
// take effect, with fresh moments and iteration count, at the first
// minimize() after they change; setMaximumIterations keeps the state.
class energyMinimizerAdam : public updater {
   public:
    explicit energyMinimizerAdam(std::shared_ptr<sphericalVertexModel> m) { setModel(std::move(m)); }
    void setAdamParameters(double b1 = 0.9, double b2 = 0.99, double eps = 0.00000001, double dt = 0.0001,
                           double fc = 1e-12) {
        p_.beta1 = b1;
        p_.beta2 = b2;
        p_.epsilon = eps;
        p_.dt = dt;
        p_.force_cutoff = fc;
        deltaT = dt;
        dirty_ = true;
    }
    void setForceCutoff(double fc) {
        p_.force_cutoff = fc;
        dirty_ = true;
    }
    void setMaximumIterations(int maxIt) override {
        maxIterations = maxIt;
        p_.max_iterations = maxIt;
        if (!dirty_) check(svm_adam_set_max_iterations(model->context(), 0, maxIt), model->context(),
                           "svm_adam_set_max_iterations");
    }
    void minimize() {
        svm_context* c = vertexModelContext("energyMinimizerAdam");
        if (dirty_) {
            p_.max_iterations = maxIterations;
            check(svm_adam_set_parameters(c, 0, &p_), c, "svm_adam_set_parameters");
            dirty_ = false;
        }
        check(svm_adam_minimize(c, 100), c, "svm_adam_minimize");
        svm_adam_state s;
        check(svm_adam_get_state(c, 0, &s), c, "svm_adam_get_state");
        iterations = s.iterations;
        forceMax = s.force_max;
    }
    void performUpdate() override { minimize(); }
    double getMaxForce() const { return forceMax; }

   private:
    svm_adam_params p_{0.9, 0.99, 1e-8, 1e-4, 1e-12, 1000, 0};
    bool dirty_ = true;
    double forceMax = 0.0;
};

// Simulation (src/simulation/simulation.cpp:102-174): updaters in order per step.
class Simulation {
   public:
    void setConfiguration(std::shared_ptr<simpleModel> c) { config = std::move(c); }
    void addUpdater(UpdaterPtr u) { updaters.push_back(std::move(u)); }
    void addUpdater(UpdaterPtr u, std::shared_ptr<simpleModel> c) {
        u->setModel(std::move(c));
        updaters.push_back(std::move(u));
    }
    void clearUpdaters() { updaters.clear(); }
    void setIntegrationTimestep(double dt) {
        for (auto& u : updaters) u->setDeltaT(dt);
    }
    void performTimestep() {
        for (auto& u : updaters) u->performUpdate();
        ++integerTimestep;
    }
    void computeForces() { config->computeForces(true); }
    double computePotentialEnergy() {
        config->computeGeometry();
        return config->computeEnergy();
    }
    double computeKineticEnergy() { return config->computeKineticEnergy(); }
    double computeEnergy() { return computePotentialEnergy() + computeKineticEnergy(); }
    int integerTimestep = 0;

   private:
    std::shared_ptr<simpleModel> config;
    std::vector<UpdaterPtr> updaters;
};

}  // namespace svm
