//! Raw bindings to the MI355X lattice-Boltzmann step engine.
//!
//! Field order and types follow `include/lbflow_hip.h` and
//! `include/lbflow_serve.h` one for one (`tests/test_rust_binding.py` checks
//! every `#[repr(C)]` struct and every function of this file against the
//! headers). All functions return 0 on success and a negative `LBF_E*` code
//! on error, with a thread-local message from `lbf_last_error`.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int};

pub const LBF_ABI_VERSION: u32 = 1;
pub const LBF_MAX_Q: usize = 27;
pub const LBF_MAX_SCALARS: usize = 4;

pub const LBF_OK: c_int = 0;
pub const LBF_EINVAL: c_int = -1;
pub const LBF_EUNSUPPORTED: c_int = -2;
pub const LBF_EDEVICE: c_int = -3;
pub const LBF_ENOMEM: c_int = -4;
pub const LBF_ESTATE: c_int = -5;
pub const LBF_ENONFINITE: c_int = -6;
pub const LBF_ETIMEOUT: c_int = -7;

// velocity sets (`VelocitySet`, src/velocity_set.rs:18-25)
pub const LBF_D2Q5: i32 = 0;
pub const LBF_D2Q9: i32 = 1;
pub const LBF_D3Q15: i32 = 2;
pub const LBF_D3Q19: i32 = 3;
pub const LBF_D3Q27: i32 = 4;

// collision operators (`CollisionOperator`, src/lib.rs:37-42)
pub const LBF_BGK: i32 = 0;
pub const LBF_TRT: i32 = 1;
pub const LBF_MRT: i32 = 2;

// faces in application order (FACES_2D / FACES_3D, src/lib.rs:27-29)
pub const LBF_WEST: usize = 0;
pub const LBF_EAST: usize = 1;
pub const LBF_SOUTH: usize = 2;
pub const LBF_NORTH: usize = 3;
pub const LBF_BOTTOM: usize = 4;
pub const LBF_TOP: usize = 5;

// momentum face conditions (src/momentum/bc.rs:11-26)
pub const LBF_BC_UNSET: i32 = 0;
pub const LBF_BC_NO_SLIP: i32 = 1;
pub const LBF_BC_BOUNCE_BACK: i32 = 2;
pub const LBF_BC_ANTI_BOUNCE_BACK: i32 = 3;
pub const LBF_BC_ZOU_HE: i32 = 4;
pub const LBF_BC_PERIODIC: i32 = 5;
pub const LBF_KNOWN_DENSITY: u32 = 1;
pub const LBF_KNOWN_UX: u32 = 2;
pub const LBF_KNOWN_UY: u32 = 4;
pub const LBF_KNOWN_UZ: u32 = 8;

// scalar face conditions (src/passive_scalar/bc.rs:9-30)
pub const LBF_SBC_UNSET: i32 = 0;
pub const LBF_SBC_ANTI_BOUNCE_BACK: i32 = 1;
pub const LBF_SBC_ANTI_BB_NO_FLUX: i32 = 2;
pub const LBF_SBC_BB_NO_FLUX: i32 = 3;
pub const LBF_SBC_ZEROTH_ORDER_NO_FLUX: i32 = 4;
pub const LBF_SBC_SECOND_ORDER_NO_FLUX: i32 = 5;
pub const LBF_SBC_PERIODIC: i32 = 6;
pub const LBF_INNER_ANTI_BOUNCE_BACK: i32 = 0;
pub const LBF_INNER_BOUNCE_BACK: i32 = 1;

// population storage
pub const LBF_F64: i32 = 0;
pub const LBF_F32: i32 = 1;
pub const LBF_F32_SHIFTED: i32 = 2;
pub const LBF_F16_SHIFTED: i32 = 3;
pub const LBF_BF16_SHIFTED: i32 = 4;

// streaming scheme
pub const LBF_STREAM_AUTO: i32 = 0;
pub const LBF_STREAM_TWO_ARRAY: i32 = 1;
pub const LBF_STREAM_AA: i32 = 2;
pub const LBF_STREAM_TWO_STEP: i32 = 3;

// sources (force / scalar source)
pub const LBF_SOURCE_NONE: i32 = 0;
pub const LBF_SOURCE_UNIFORM: i32 = 1;
pub const LBF_SOURCE_FIELD: i32 = 2;

// case and step flags
pub const LBF_CASE_FREEZE_MOMENTUM: u32 = 1;
pub const LBF_CASE_ZOU_HE_3D: u32 = 2;
pub const LBF_STEP_RESIDUALS: u32 = 1;
pub const LBF_STEP_STOP_ON_CONVERGENCE: u32 = 2;

// reductions
pub const LBF_REDUCE_MEAN_DENSITY: i32 = 0;
pub const LBF_REDUCE_TOTAL_MASS: i32 = 1;
pub const LBF_REDUCE_SCALAR_TOTAL: i32 = 2;
pub const LBF_REDUCE_ADSORBED: i32 = 3;
pub const LBF_REDUCE_NONFINITE: i32 = 4;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct lbf_face_bc {
    pub kind: i32,
    pub known: u32,
    pub density: f64,
    pub velocity: [f64; 3],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct lbf_collision {
    pub op: i32,
    pub reserved: i32,
    pub tau: f64,
    pub omega_plus: f64,
    pub omega_minus: f64,
    pub relaxation: [f64; LBF_MAX_Q],
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct lbf_scalar_face_bc {
    pub kind: i32,
    pub unknown_populations_only: i32,
    pub scalar_value: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct lbf_scalar_desc {
    pub collision: lbf_collision,
    pub velocity_set: i32,
    pub inner_bc: i32,
    pub faces: [lbf_scalar_face_bc; 6],
    pub initial_value: *const f64,
    pub initial_value_uniform: f64,
    pub source_kind: i32,
    pub has_adsorption: i32,
    pub source_uniform: f64,
    pub source_field: *const f64,
    pub adsorption_rate: f64,
    pub desorption_rate: f64,
    pub q_max: f64,
    pub initial_q_ads: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct lbf_case_desc {
    pub abi_version: u32,
    pub velocity_set: i32,
    pub n: [u32; 3],
    pub flags: u32,
    pub collision: lbf_collision,
    pub node_type: *const u8,
    pub faces: [lbf_face_bc; 6],
    pub force_kind: i32,
    pub precision: i32,
    pub force: [f64; 3],
    pub force_field: *const f64,
    pub initial_density: *const f64,
    pub initial_density_uniform: f64,
    pub initial_velocity: *const f64,
    pub initial_velocity_uniform: [f64; 3],
    pub n_scalars: u32,
    pub streaming: i32,
    pub scalars: *const lbf_scalar_desc,
    pub shift_scale: f64,
    pub multiphase: i32,
    pub reserved: i32,
    pub g_function: f64,
    pub wall_density: f64,
}

/// Opaque case handle (device buffers and one HIP stream).
#[repr(C)]
pub struct lbf_case {
    _private: [u8; 0],
}

extern "C" {
    pub fn lbf_create(desc: *const lbf_case_desc, out: *mut *mut lbf_case) -> c_int;
    pub fn lbf_destroy(c: *mut lbf_case);
    pub fn lbf_initialize_nodes(c: *mut lbf_case) -> c_int;
    pub fn lbf_main_steps(c: *mut lbf_case, nsteps: u32, flags: u32) -> c_int;
    pub fn lbf_synchronize(c: *mut lbf_case) -> c_int;
    pub fn lbf_run(c: *mut lbf_case, max_iterations: u64, steps_done: *mut u64) -> c_int;
    pub fn lbf_residuals(c: *mut lbf_case, r_rho: *mut f64, r_u: *mut f64, r_c: *mut f64) -> c_int;
    pub fn lbf_time_step(c: *mut lbf_case, t: *mut u64) -> c_int;
    pub fn lbf_reduce(c: *mut lbf_case, which: i32, out: *mut f64) -> c_int;
    pub fn lbf_get_fields(c: *mut lbf_case, rho: *mut f64, u: *mut f64) -> c_int;
    pub fn lbf_get_scalar(c: *mut lbf_case, s: u32, value: *mut f64, q_ads: *mut f64) -> c_int;
    pub fn lbf_get_pops(c: *mut lbf_case, field: i32, out: *mut f64) -> c_int;
    pub fn lbf_get_pops_noalloc(c: *mut lbf_case, field: i32, out: *mut f64) -> c_int;
    pub fn lbf_set_pops(c: *mut lbf_case, field: i32, input: *const f64) -> c_int;
    pub fn lbf_get_link_tables(
        c: *mut lbf_case,
        node_type: *mut u8,
        bb_link_mask: *mut u32,
        face_bits: *mut u8,
    ) -> c_int;
    pub fn lbf_device_bytes(desc: *const lbf_case_desc, bytes: *mut u64) -> c_int;
    pub fn lbf_last_error() -> *const c_char;
}

// ---- serving tier (include/lbflow_serve.h) ------------------------------

pub const LBF_SERVE_KEEP_POPS: u32 = 1;
pub const LBF_REQ_QUEUED: i32 = 0;
pub const LBF_REQ_PREPARING: i32 = 1;
pub const LBF_REQ_RUNNING: i32 = 2;
pub const LBF_REQ_DONE: i32 = 3;
pub const LBF_REQ_FAILED: i32 = 4;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct lbf_server_config {
    pub abi_version: u32,
    pub max_slots: u32,
    pub hbm_budget_bytes: u64,
    pub max_steps_per_round: u32,
    pub autotune: i32,
    pub grid: u32,
    pub tile: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct lbf_request_info {
    pub state: i32,
    pub error: i32,
    pub nodes: u64,
    pub steps: u64,
    pub device_bytes: u64,
    pub t_submit: f64,
    pub t_admit: f64,
    pub t_start: f64,
    pub t_done: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct lbf_server_stats {
    pub submitted: u64,
    pub completed: u64,
    pub failed: u64,
    pub queued: u64,
    pub resident: u64,
    pub rounds: u64,
    pub graphs_built: u64,
    pub hbm_in_use: u64,
    pub hbm_peak: u64,
    pub hbm_budget: u64,
    pub node_updates: f64,
    pub first_submit: f64,
    pub last_done: f64,
    pub gpu_seconds: f64,
    pub p50_latency: f64,
    pub p99_latency: f64,
    pub mean_latency: f64,
    pub cases_per_second: f64,
    pub mlups: f64,
    pub prep_seconds: f64,
    pub retire_seconds: f64,
    pub tuned_grid: u32,
    pub tuned_tile: u32,
}

/// Opaque server handle (worker and helper threads, resident cases).
#[repr(C)]
pub struct lbf_server {
    _private: [u8; 0],
}

extern "C" {
    pub fn lbf_server_create(cfg: *const lbf_server_config, out: *mut *mut lbf_server) -> c_int;
    pub fn lbf_server_destroy(s: *mut lbf_server);
    pub fn lbf_server_submit(
        s: *mut lbf_server,
        desc: *const lbf_case_desc,
        steps: u64,
        flags: u32,
        ticket: *mut u64,
    ) -> c_int;
    pub fn lbf_server_info(s: *mut lbf_server, ticket: u64, info: *mut lbf_request_info) -> c_int;
    pub fn lbf_server_wait(s: *mut lbf_server, ticket: u64, timeout_s: f64) -> c_int;
    pub fn lbf_server_result(
        s: *mut lbf_server,
        ticket: u64,
        rho: *mut f64,
        u: *mut f64,
        pops: *mut f64,
        release: i32,
    ) -> c_int;
    pub fn lbf_server_drain(s: *mut lbf_server, timeout_s: f64) -> c_int;
    pub fn lbf_server_get_stats(s: *mut lbf_server, out: *mut lbf_server_stats) -> c_int;
    pub fn lbf_server_now(s: *mut lbf_server) -> f64;
}
