// diff.cpp -- host side of differentiable rollouts (see diff.hpp): flattening
// of `Physics` worlds, device buffers, checkpointed forward and the
// segment-by-segment adjoint.
#include "diff.hpp"

#include <cmath>

#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <string>
#include <vector>

#include "world.hpp"

namespace bg {

#define HIP_TRY(expr)                                                                                      \
    do {                                                                                                   \
        hipError_t _e = (expr);                                                                            \
        if (_e != hipSuccess) return fail(BG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

namespace {

template <class T>
hipError_t dalloc(T** p, size_t n) {
    *p = nullptr;
    return hipMalloc((void**)p, (n ? n : 1) * sizeof(T));
}
template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}
template <class T>
hipError_t up(T* dst, const std::vector<T>& v, hipStream_t s) {
    return v.empty() ? hipSuccess : hipMemcpyAsync(dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

}  // namespace

int World::export_diff(DiffHostWorld& out) {
    if (sync_host()) return fail(BG_ERR_HIP, "download failed");
    if (!constraints_.empty())
        return fail(BG_ERR_UNSUPPORTED, "differentiable rollouts take worlds without circle constraints");
    if (params_.collisions_enabled && params_.use_spatial_hash)
        return fail(BG_ERR_UNSUPPORTED, "spatial collisions not supported right now (physics.rs:411-412)");
    std::vector<uint32_t> bs, cs;
    dense_slots(bs, cs);
    const uint32_t nb = (uint32_t)bs.size();
    if (nb > kDiffMaxBodies)
        return fail(BG_ERR_CAPACITY, "a differentiable world holds at most " + std::to_string(kDiffMaxBodies) +
                                         " bodies");
    if (cs.size() != nb) return fail(BG_ERR_UNSUPPORTED, "differentiable rollouts take one sphere collider per body");
    out = DiffHostWorld{};
    out.pos.resize(nb);
    out.pos_old.resize(nb);
    out.vel.resize(nb);
    out.vreq.resize(nb);
    out.acc.resize(nb);
    out.mass.resize(nb);
    out.radius.resize(nb);
    out.gmod.resize(nb);
    out.flags.resize(nb);
    std::vector<int32_t> dense_of_slot(bodies_.capacity_slots(), -1);
    for (uint32_t k = 0; k < nb; ++k) dense_of_slot[bs[k]] = (int32_t)k;
    for (uint32_t k = 0; k < nb; ++k) {
        const HostBody& hb = bodies_.at_slot(bs[k]);
        const bg_rigid_body& f = hb.f;
        const bg_collider& c = colliders_.at_slot(cs[k]).f;
        const bg_handle ch = colliders_.handle_at(cs[k]);
        bool ok = c.has_parent && c.parent == bodies_.handle_at(bs[k]) && !hb.colliders.empty();
        for (bg_handle h : hb.colliders) ok = ok && h == ch;
        if (!ok)
            return fail(BG_ERR_UNSUPPORTED,
                        "differentiable rollouts take one sphere collider per body, inserted in body order");
        if (c.offset_translation[0] != 0.0f || c.offset_translation[1] != 0.0f || c.is_sensor ||
            c.memberships != 0xFFFFFFFFu || c.filter != 0xFFFFFFFFu)
            return fail(BG_ERR_UNSUPPORTED,
                        "differentiable rollouts take centred, non-sensor colliders with default interaction groups");
        out.pos[k] = make_float2(f.position[0], f.position[1]);
        out.pos_old[k] = make_float2(f.position_old[0], f.position_old[1]);
        out.vel[k] = make_float2(f.calculated_velocity[0], f.calculated_velocity[1]);
        out.vreq[k] = make_float2(f.velocity_request[0], f.velocity_request[1]);
        out.acc[k] = make_float2(f.acceleration[0], f.acceleration[1]);
        out.mass[k] = f.calculated_mass;
        out.radius[k] = c.radius;
        out.gmod[k] = f.gravity_mod;
        uint32_t fl = 0;
        if (f.body_type == BG_STATIC) fl |= BF_STATIC;
        if (f.has_velocity_request) fl |= BF_HAS_VREQ;
        out.flags[k] = fl;
        // the collider snapshot must be the body position (zero offset, any rotation)
        if (c.absolute_translation[0] != f.position[0] || c.absolute_translation[1] != f.position[1])
            return fail(BG_ERR_UNSUPPORTED, "collider absolute translation must equal its body position at the start");
    }
    for (uint32_t s = 0; s < springs_.capacity_slots(); ++s) {
        if (!springs_.occupied(s)) continue;
        const bg_spring& sp = springs_.at_slot(s);
        const HostBody* a = bodies_.get(sp.rigid_body_a);
        const HostBody* b = bodies_.get(sp.rigid_body_b);
        if (!a || !b || sp.rigid_body_a == sp.rigid_body_b)
            return fail(BG_ERR_PANIC, "spring endpoints must be two existing, distinct bodies (springs.rs:26-29 unwrap)");
        out.sa.push_back(dense_of_slot[Arena<HostBody>::slot_of(sp.rigid_body_a)]);
        out.sb.push_back(dense_of_slot[Arena<HostBody>::slot_of(sp.rigid_body_b)]);
        out.rest.push_back(sp.rest_length);
        out.k.push_back(sp.stiffness);
        out.c.push_back(sp.damping);
    }
    for (uint32_t s = 0; s < joints_.capacity_slots(); ++s) {  // arena order (physics.rs:426)
        if (!joints_.occupied(s)) continue;
        const bg_fixed_joint& j = joints_.at_slot(s);
        const HostBody* a = bodies_.get(j.rigid_body_a);
        const HostBody* b = bodies_.get(j.rigid_body_b);
        if (!a || !b || j.rigid_body_a == j.rigid_body_b)
            return fail(BG_ERR_PANIC, "joint bodies must exist and differ (physics.rs:427-432 unwrap)");
        if (!(a->f.calculated_mass > 0.0f) || !(b->f.calculated_mass > 0.0f))
            return fail(BG_ERR_PANIC, "joint bodies need a positive mass (physics.rs:447-448 assert)");
        out.ja.push_back(dense_of_slot[Arena<HostBody>::slot_of(j.rigid_body_a)]);
        out.jb.push_back(dense_of_slot[Arena<HostBody>::slot_of(j.rigid_body_b)]);
        out.jan_a.push_back(make_float2(j.anchor_a[0], j.anchor_a[1]));
        out.jan_b.push_back(make_float2(j.anchor_b[0], j.anchor_b[1]));
        out.jdist.push_back(j.distance);
    }
    out.joint_iterations = params_.joint_iterations;
    out.old_dt = params_.old_dt;
    out.gx = params_.gravity[0];
    out.gy = params_.gravity[1];
    out.substeps = params_.substeps;
    out.collisions = params_.collisions_enabled != 0;
    return BG_OK;
}

Rollout::~Rollout() { release(); }

void Rollout::release() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    dfree(d_worlds_), dfree(d_mass_), dfree(d_radius_), dfree(d_gmod_), dfree(d_flags_), dfree(d_vreq_),
        dfree(d_acc0_);
    dfree(d_sa_), dfree(d_sb_), dfree(d_rest_), dfree(d_k_), dfree(d_c_), dfree(d_inc_start_), dfree(d_inc_list_);
    dfree(d_ck_p_), dfree(d_ck_q_), dfree(d_ck_v_), dfree(d_final_), dfree(d_loss_);
    dfree(d_tape_s_), dfree(d_tape_nbr_), dfree(d_lam_p_), dfree(d_lam_q_), dfree(d_dl_), dfree(d_gv_),
        dfree(d_gk_);
    dfree(d_tape_v_), dfree(d_tape_pj_), dfree(d_j_axis_), dfree(d_lam_v_), dfree(d_jc_start_), dfree(d_j_ab_),
        dfree(d_j_anchor_), dfree(d_j_mass_), dfree(d_j_dist_);
    if (stream_) (void)hipStreamDestroy(stream_);
    stream_ = nullptr;
    n_ck_alloc_ = tape_sub_alloc_ = 0;
}

int Rollout::create(World* const* worlds, uint32_t n, const float* targets, uint32_t checkpoint_every) {
    if (n == 0 || !worlds || !targets) return fail(BG_ERR_INVALID_ARGUMENT, "worlds, n > 0 and targets are required");
    if (checkpoint_every == 0) return fail(BG_ERR_INVALID_ARGUMENT, "checkpoint_every must be >= 1");
    int dev_n = 0;
    if (hipGetDeviceCount(&dev_n) != hipSuccess || dev_n == 0)
        return fail(BG_ERR_NO_DEVICE, "no HIP device available: the rollout has no CPU fallback");
    ck_every_ = checkpoint_every;
    std::vector<DiffWorld> ws(n);
    std::vector<float2> p0, q0, v0;
    std::vector<uint32_t> inc_count;
    std::vector<uint32_t> jc_start(1, 0), j_ab;
    std::vector<float4> j_anchor, j_mass;
    std::vector<float> j_dist;
    damped_ = false;
    for (uint32_t w = 0; w < n; ++w) {
        if (!worlds[w]) return fail(BG_ERR_INVALID_ARGUMENT, "null world");
        DiffHostWorld h;
        int st = worlds[w]->export_diff(h);
        if (st) return fail(st, "world " + std::to_string(w) + ": " + worlds[w]->last_error);
        if (w == 0) {
            gx_ = h.gx;
            gy_ = h.gy;
            substeps_ = h.substeps;
            collisions_ = h.collisions;
            j_iters_ = h.joint_iterations;
        } else if (h.gx != gx_ || h.gy != gy_ || h.substeps != substeps_ || h.collisions != collisions_ ||
                   h.joint_iterations != j_iters_) {
            return fail(BG_ERR_INVALID_ARGUMENT,
                        "worlds of a rollout share gravity, substeps, joint_iterations and collisions_enabled");
        }
        DiffWorld& d = ws[w];
        d.b0 = (uint32_t)mass_.size();
        d.nb = (uint32_t)h.mass.size();
        d.s0 = (uint32_t)k_.size();
        d.ns = (uint32_t)h.k.size();
        d.old_dt0 = h.old_dt;
        d.first_dyn = -1;
        d.rmax = 0.0f;
        d.tx = targets[2 * w];
        d.ty = targets[2 * w + 1];
        for (uint32_t i = 0; i < d.nb; ++i) {
            if (d.first_dyn < 0 && !(h.flags[i] & BF_STATIC)) d.first_dyn = (int32_t)i;
            if (!(h.radius[i] <= d.rmax)) d.rmax = h.radius[i];
        }
        p0.insert(p0.end(), h.pos.begin(), h.pos.end());
        q0.insert(q0.end(), h.pos_old.begin(), h.pos_old.end());
        v0.insert(v0.end(), h.vel.begin(), h.vel.end());
        vreq_.insert(vreq_.end(), h.vreq.begin(), h.vreq.end());
        acc0_.insert(acc0_.end(), h.acc.begin(), h.acc.end());
        mass_.insert(mass_.end(), h.mass.begin(), h.mass.end());
        radius_.insert(radius_.end(), h.radius.begin(), h.radius.end());
        gmod_.insert(gmod_.end(), h.gmod.begin(), h.gmod.end());
        flags_.insert(flags_.end(), h.flags.begin(), h.flags.end());
        // joints: connected components of the world's joint graph (union-find over bodies), each
        // component's joints contiguous and in arena order; components touch disjoint bodies, so
        // sweeping them concurrently is the reference's sequential sweep
        d.jc0 = (uint32_t)jc_start.size() - 1;
        d.njc = 0;
        if (!h.ja.empty()) {
            std::vector<int32_t> uf(d.nb);
            for (uint32_t i = 0; i < d.nb; ++i) uf[i] = (int32_t)i;
            auto find = [&](int32_t x) {
                while (uf[x] != x) x = uf[x] = uf[uf[x]];
                return x;
            };
            for (size_t q = 0; q < h.ja.size(); ++q) {
                const int32_t ra = find(h.ja[q]), rb = find(h.jb[q]);
                if (ra != rb) uf[ra] = rb;
            }
            std::vector<int32_t> comp_of_root(d.nb, -1);
            std::vector<std::vector<uint32_t>> comps;
            for (size_t q = 0; q < h.ja.size(); ++q) {
                const int32_t r = find(h.ja[q]);
                if (comp_of_root[r] < 0) {
                    comp_of_root[r] = (int32_t)comps.size();
                    comps.emplace_back();
                }
                comps[comp_of_root[r]].push_back((uint32_t)q);
            }
            for (const auto& cq : comps) {
                for (uint32_t q : cq) {
                    const float ma = h.mass[h.ja[q]], mb = h.mass[h.jb[q]];
                    const float inv_a = 1.0f / ma, inv_b = 1.0f / mb;  // physics.rs:450-455
                    const float ims = inv_a + inv_b;
                    const float ratio = inv_a / ims;
                    j_ab.push_back((uint32_t)h.ja[q] | ((uint32_t)h.jb[q] << 16));
                    j_anchor.push_back(make_float4(h.jan_a[q].x, h.jan_a[q].y, h.jan_b[q].x, h.jan_b[q].y));
                    j_dist.push_back(h.jdist[q]);
                    j_mass.push_back(make_float4(ims, ratio, 1.0f - ratio, 1.0f));
                }
                jc_start.push_back((uint32_t)j_ab.size());
            }
            d.njc = (uint32_t)comps.size();
        }
        for (size_t s = 0; s < h.k.size(); ++s) {
            sa_.push_back((int32_t)d.b0 + h.sa[s]);
            sb_.push_back((int32_t)d.b0 + h.sb[s]);
            rest_.push_back(h.rest[s]);
            k_.push_back(h.k[s]);
            c_.push_back(h.c[s]);
            if (h.c[s] != 0.0f) damped_ = true;
        }
    }
    nb_ = (uint32_t)mass_.size();
    ns_ = (uint32_t)k_.size();
    nj_ = (uint32_t)j_ab.size();
    n_worlds_ = n;
    body0_.resize(n + 1);
    for (uint32_t w = 0; w < n; ++w) body0_[w] = ws[w].b0;
    body0_[n] = nb_;
    spring_world_.resize(ns_);
    for (uint32_t w = 0; w < n; ++w)
        for (uint32_t s = ws[w].s0; s < ws[w].s0 + ws[w].ns; ++s) spring_world_[s] = w;
    nonfinite_.assign(n, 0);
    // per body: its springs in ascending index (summation order of physics.rs:413 + springs.rs:25-47)
    std::vector<uint32_t> inc_start(nb_ + 1, 0), inc_list(2 * (size_t)ns_);
    for (uint32_t s = 0; s < ns_; ++s) {
        inc_start[sa_[s] + 1]++;
        inc_start[sb_[s] + 1]++;
    }
    for (uint32_t i = 0; i < nb_; ++i) inc_start[i + 1] += inc_start[i];
    std::vector<uint32_t> fill(inc_start.begin(), inc_start.end() - 1);
    for (uint32_t s = 0; s < ns_; ++s) {
        inc_list[fill[sa_[s]]++] = s;
        inc_list[fill[sb_[s]]++] = s;
    }
    HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    hipStream_t s = stream_;
    HIP_TRY(dalloc(&d_worlds_, n));
    HIP_TRY(dalloc(&d_mass_, nb_));
    HIP_TRY(dalloc(&d_radius_, nb_));
    HIP_TRY(dalloc(&d_gmod_, nb_));
    HIP_TRY(dalloc(&d_flags_, nb_));
    HIP_TRY(dalloc(&d_vreq_, nb_));
    HIP_TRY(dalloc(&d_acc0_, nb_));
    HIP_TRY(dalloc(&d_sa_, ns_));
    HIP_TRY(dalloc(&d_sb_, ns_));
    HIP_TRY(dalloc(&d_rest_, ns_));
    HIP_TRY(dalloc(&d_k_, ns_));
    HIP_TRY(dalloc(&d_c_, ns_));
    HIP_TRY(dalloc(&d_inc_start_, nb_ + 1));
    HIP_TRY(dalloc(&d_inc_list_, inc_list.size()));
    HIP_TRY(dalloc(&d_final_, nb_));
    HIP_TRY(dalloc(&d_loss_, n));
    HIP_TRY(dalloc(&d_lam_p_, nb_));
    HIP_TRY(dalloc(&d_lam_q_, nb_));
    HIP_TRY(dalloc(&d_dl_, nb_));
    HIP_TRY(dalloc(&d_gv_, nb_));
    HIP_TRY(dalloc(&d_gk_, ns_));
    HIP_TRY(dalloc(&d_lam_v_, nb_));
    HIP_TRY(dalloc(&d_jc_start_, jc_start.size()));
    HIP_TRY(dalloc(&d_j_ab_, nj_));
    HIP_TRY(dalloc(&d_j_anchor_, nj_));
    HIP_TRY(dalloc(&d_j_mass_, nj_));
    HIP_TRY(dalloc(&d_j_dist_, nj_));
    HIP_TRY(dalloc(&d_j_axis_, (size_t)nj_ * (j_iters_ ? j_iters_ : 1)));
    HIP_TRY(up(d_jc_start_, jc_start, s));
    HIP_TRY(up(d_j_ab_, j_ab, s));
    HIP_TRY(up(d_j_anchor_, j_anchor, s));
    HIP_TRY(up(d_j_mass_, j_mass, s));
    HIP_TRY(up(d_j_dist_, j_dist, s));
    HIP_TRY(up(d_worlds_, ws, s));
    HIP_TRY(up(d_mass_, mass_, s));
    HIP_TRY(up(d_radius_, radius_, s));
    HIP_TRY(up(d_gmod_, gmod_, s));
    HIP_TRY(up(d_flags_, flags_, s));
    HIP_TRY(up(d_vreq_, vreq_, s));
    HIP_TRY(up(d_acc0_, acc0_, s));
    HIP_TRY(up(d_sa_, sa_, s));
    HIP_TRY(up(d_sb_, sb_, s));
    HIP_TRY(up(d_rest_, rest_, s));
    HIP_TRY(up(d_k_, k_, s));
    HIP_TRY(up(d_c_, c_, s));
    HIP_TRY(up(d_inc_start_, inc_start, s));
    HIP_TRY(up(d_inc_list_, inc_list, s));
    p0_ = p0;
    q0_ = q0;
    v0_ = v0;
    HIP_TRY(hipStreamSynchronize(s));
    return BG_OK;
}

int Rollout::set_params(const float* vreq, const float* k) {
    if (vreq) {
        memcpy(vreq_.data(), vreq, sizeof(float2) * nb_);
        HIP_TRY(up(d_vreq_, vreq_, stream_));
    }
    if (k) {
        memcpy(k_.data(), k, sizeof(float) * ns_);
        HIP_TRY(up(d_k_, k_, stream_));
    }
    HIP_TRY(hipStreamSynchronize(stream_));
    have_forward_ = false;
    return BG_OK;
}

int Rollout::ensure_buffers(uint32_t n_steps) {
    const uint32_t n_ck = (n_steps + ck_every_ - 1) / ck_every_;
    const uint32_t n_ck_buf = n_ck > 0 ? n_ck : 1;
    if (n_ck_buf > n_ck_alloc_) {
        dfree(d_ck_p_), dfree(d_ck_q_), dfree(d_ck_v_);
        HIP_TRY(dalloc(&d_ck_p_, (size_t)n_ck_buf * nb_));
        HIP_TRY(dalloc(&d_ck_q_, (size_t)n_ck_buf * nb_));
        HIP_TRY(dalloc(&d_ck_v_, (size_t)n_ck_buf * nb_));
        n_ck_alloc_ = n_ck_buf;
    }
    const uint32_t seg_sub = (n_steps < ck_every_ ? n_steps : ck_every_) * substeps_;
    if (seg_sub > tape_sub_alloc_) {
        dfree(d_tape_s_), dfree(d_tape_nbr_), dfree(d_tape_v_), dfree(d_tape_pj_);
        HIP_TRY(dalloc(&d_tape_s_, (size_t)seg_sub * nb_));
        HIP_TRY(dalloc(&d_tape_nbr_, (size_t)seg_sub * nb_ * kDiffTapeNbr));
        if (damped_) HIP_TRY(dalloc(&d_tape_v_, (size_t)seg_sub * nb_));
        if (nj_ > 0 && j_iters_ > 0) HIP_TRY(dalloc(&d_tape_pj_, (size_t)seg_sub * nb_));
        tape_sub_alloc_ = seg_sub;
    }
    return BG_OK;
}

DiffArgs Rollout::args() const {
    DiffArgs a;
    memset(&a, 0, sizeof(a));
    a.worlds = d_worlds_;
    a.n_worlds = n_worlds_;
    a.nb_total = nb_;
    a.mass = d_mass_;
    a.radius = d_radius_;
    a.gmod = d_gmod_;
    a.flags = d_flags_;
    a.vreq = d_vreq_;
    a.acc0 = d_acc0_;
    a.sp_a = d_sa_;
    a.sp_b = d_sb_;
    a.sp_rest = d_rest_;
    a.sp_k = d_k_;
    a.sp_c = d_c_;
    a.inc_start = d_inc_start_;
    a.inc_list = d_inc_list_;
    a.jc_start = d_jc_start_;
    a.j_ab = d_j_ab_;
    a.j_anchor = d_j_anchor_;
    a.j_dist = d_j_dist_;
    a.j_mass = d_j_mass_;
    a.n_joints = nj_;
    a.j_iters = nj_ > 0 ? j_iters_ : 0u;
    a.ck_p = d_ck_p_;
    a.ck_q = d_ck_q_;
    a.ck_v = d_ck_v_;
    a.final_p = d_final_;
    a.loss = d_loss_;
    a.tape_s = d_tape_s_;
    a.tape_nbr = d_tape_nbr_;
    a.tape_v = d_tape_v_;
    a.tape_pj = d_tape_pj_;
    a.j_axis = d_j_axis_;
    a.lam_p = d_lam_p_;
    a.lam_q = d_lam_q_;
    a.lam_v = d_lam_v_;
    a.dl_dfinal = nullptr;
    a.grad_vreq = d_gv_;
    a.grad_k = d_gk_;
    a.gx = gx_;
    a.gy = gy_;
    a.h = h_;
    a.substeps = substeps_;
    a.collisions = collisions_ ? 1u : 0u;
    return a;
}

int Rollout::forward(double delta, uint32_t n_steps, float* loss_per_world) {
    if (n_steps == 0) return fail(BG_ERR_INVALID_ARGUMENT, "n_steps must be >= 1");
    int st = ensure_buffers(n_steps);
    if (st) return st;
    h_ = (float)delta / (float)substeps_;  // physics.rs:80, 399
    // checkpoint 0 = the initial state
    HIP_TRY(up(d_ck_p_, p0_, stream_));
    HIP_TRY(up(d_ck_q_, q0_, stream_));
    HIP_TRY(up(d_ck_v_, v0_, stream_));
    HIP_TRY(launch_diff_forward(args(), n_steps, ck_every_, stream_));
    std::vector<float> loss(n_worlds_);
    HIP_TRY(hipMemcpyAsync(loss.data(), d_loss_, sizeof(float) * n_worlds_, hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    if (loss_per_world) memcpy(loss_per_world, loss.data(), sizeof(float) * n_worlds_);
    n_steps_ = n_steps;
    have_forward_ = true;
    return BG_OK;
}

int Rollout::backward(const float* dl_dfinal, float* grad_vreq, float* grad_k) {
    if (!have_forward_) return fail(BG_ERR_INVALID_ARGUMENT, "backward needs a forward with the current parameters");
    DiffArgs a = args();
    if (dl_dfinal) {
        std::vector<float2> dl(nb_);
        memcpy(dl.data(), dl_dfinal, sizeof(float2) * nb_);
        HIP_TRY(up(d_dl_, dl, stream_));
        a.dl_dfinal = d_dl_;
    }
    HIP_TRY(hipMemsetAsync(d_gk_, 0, sizeof(float) * (ns_ ? ns_ : 1), stream_));
    const uint32_t n_seg = (n_steps_ + ck_every_ - 1) / ck_every_;
    for (uint32_t seg = n_seg; seg-- > 0;) {
        const uint32_t first = seg * ck_every_;
        const uint32_t steps = (n_steps_ - first) < ck_every_ ? (n_steps_ - first) : ck_every_;
        HIP_TRY(launch_diff_backward_segment(a, seg, first, steps, seg == n_seg - 1, stream_));
    }
    std::vector<float2> gv(nb_);
    std::vector<float> gk(ns_);
    HIP_TRY(hipMemcpyAsync(gv.data(), d_gv_, sizeof(float2) * nb_, hipMemcpyDeviceToHost, stream_));
    if (ns_) HIP_TRY(hipMemcpyAsync(gk.data(), d_gk_, sizeof(float) * ns_, hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    // worlds whose gradients left the f32 range (reported, not an error: the fixed-pair
    // derivative of the contact projection grows with the horizon)
    nonfinite_.assign(n_worlds_, 0);
    for (uint32_t w = 0; w < n_worlds_; ++w)
        for (uint32_t i = body0_[w]; i < body0_[w + 1] && !nonfinite_[w]; ++i)
            if (!std::isfinite(gv[i].x) || !std::isfinite(gv[i].y)) nonfinite_[w] = 1;
    for (uint32_t s = 0; s < ns_; ++s)
        if (!std::isfinite(gk[s])) nonfinite_[spring_world_[s]] = 1;
    if (grad_vreq) memcpy(grad_vreq, gv.data(), sizeof(float2) * nb_);
    if (grad_k && ns_) memcpy(grad_k, gk.data(), sizeof(float) * ns_);
    return BG_OK;
}

int Rollout::nonfinite(uint8_t* per_world) const {
    if (!per_world) return BG_ERR_INVALID_ARGUMENT;
    for (uint32_t w = 0; w < n_worlds_; ++w) per_world[w] = w < nonfinite_.size() ? nonfinite_[w] : 0;
    return BG_OK;
}

int Rollout::final_positions(float* out) {
    if (!have_forward_) return fail(BG_ERR_INVALID_ARGUMENT, "no forward has run");
    HIP_TRY(hipMemcpyAsync(out, d_final_, sizeof(float2) * nb_, hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    return BG_OK;
}

}  // namespace bg
