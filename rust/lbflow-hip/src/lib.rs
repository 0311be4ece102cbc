//! A lattice case resident on one MI355X, driven with the method names of
//! lbflow's `Lattice` (`initialize_nodes`, `main_steps`, the time loop with
//! its stop rule, residuals, mean density, fields) so that the time loops of
//! src/momentum.rs:273-292 and src/passive_scalar.rs:151-176 can call it in
//! place of the per-node CPU update. Errors follow lbflow's
//! `LbResult<T> = Result<T, Box<dyn Error>>` (src/cli.rs:5) and carry
//! `lbf_last_error()`.
use std::error::Error;
use std::ffi::CStr;
use std::fmt;
use std::ptr;

pub use lbflow_hip_sys as sys;

pub type LbResult<T> = Result<T, Box<dyn Error>>;

#[derive(Debug)]
pub struct HipError {
    pub code: i32,
    pub message: String,
}

impl fmt::Display for HipError {
    fn fmt(&self, f: &mut fmt::Formatter<'_>) -> fmt::Result {
        write!(f, "lbflow_hip error {}: {}", self.code, self.message)
    }
}

impl Error for HipError {}

fn check(code: i32) -> LbResult<()> {
    if code == sys::LBF_OK {
        return Ok(());
    }
    // SAFETY: lbf_last_error returns a NUL-terminated thread-local string.
    let message = unsafe { CStr::from_ptr(sys::lbf_last_error()) }.to_string_lossy().into_owned();
    Err(Box::new(HipError { code, message }))
}

/// Collision operator (`CollisionOperator`, src/lib.rs:37-42).
#[derive(Clone, Debug)]
pub enum Collision {
    Bgk { tau: f64 },
    Trt { omega_plus: f64, omega_minus: f64 },
    Mrt { relaxation: Vec<f64> },
}

impl Collision {
    fn to_c(&self) -> sys::lbf_collision {
        let mut c = sys::lbf_collision {
            op: sys::LBF_BGK,
            reserved: 0,
            tau: 0.0,
            omega_plus: 0.0,
            omega_minus: 0.0,
            relaxation: [0.0; sys::LBF_MAX_Q],
        };
        match self {
            Collision::Bgk { tau } => c.tau = *tau,
            Collision::Trt { omega_plus, omega_minus } => {
                c.op = sys::LBF_TRT;
                c.omega_plus = *omega_plus;
                c.omega_minus = *omega_minus;
            }
            Collision::Mrt { relaxation } => {
                c.op = sys::LBF_MRT;
                c.relaxation[..relaxation.len()].copy_from_slice(relaxation);
            }
        }
        c
    }
}

/// Momentum face condition (`momentum::bc::BoundaryCondition`, src/momentum/bc.rs:11-26).
#[derive(Clone, Copy, Debug)]
pub enum FaceBc {
    NoSlip,
    BounceBack { density: f64, velocity: [f64; 3] },
    AntiBounceBack { density: f64 },
    /// Zou-He: `None` marks an unknown component (Option<Float> in the reference).
    ZouHe { density: Option<f64>, velocity: [Option<f64>; 3] },
    Periodic,
}

impl FaceBc {
    fn to_c(self) -> sys::lbf_face_bc {
        let mut f = sys::lbf_face_bc::default();
        match self {
            FaceBc::NoSlip => f.kind = sys::LBF_BC_NO_SLIP,
            FaceBc::BounceBack { density, velocity } => {
                f.kind = sys::LBF_BC_BOUNCE_BACK;
                f.density = density;
                f.velocity = velocity;
            }
            FaceBc::AntiBounceBack { density } => {
                f.kind = sys::LBF_BC_ANTI_BOUNCE_BACK;
                f.density = density;
            }
            FaceBc::ZouHe { density, velocity } => {
                f.kind = sys::LBF_BC_ZOU_HE;
                if let Some(d) = density {
                    f.known |= sys::LBF_KNOWN_DENSITY;
                    f.density = d;
                }
                let bits = [sys::LBF_KNOWN_UX, sys::LBF_KNOWN_UY, sys::LBF_KNOWN_UZ];
                for (a, v) in velocity.iter().enumerate() {
                    if let Some(v) = v {
                        f.known |= bits[a];
                        f.velocity[a] = *v;
                    }
                }
            }
            FaceBc::Periodic => f.kind = sys::LBF_BC_PERIODIC,
        }
        f
    }
}

/// A momentum-only case description. Arrays are in the reference's node
/// order x + nx (y + ny z) (src/momentum/lattice.rs:70-85); the reference's
/// force closure is evaluated once per node into `force_field` by the
/// caller, or given as one uniform vector.
#[derive(Clone, Debug)]
pub struct CaseSpec {
    pub velocity_set: i32,
    pub n: [u32; 3],
    pub collision: Collision,
    pub faces: [FaceBc; 6],
    pub node_type: Option<Vec<u8>>,
    pub force: Option<[f64; 3]>,
    pub force_field: Option<Vec<f64>>,
    pub initial_density: f64,
    pub initial_velocity: [f64; 3],
    pub precision: i32,
    pub streaming: i32,
    pub flags: u32,
}

/// One case on the device: the replacement of lbflow's momentum `Lattice`.
pub struct HipLattice {
    handle: *mut sys::lbf_case,
    d: usize,
    n_nodes: usize,
    q: usize,
}

// The handle is owned and driven from one thread at a time.
unsafe impl Send for HipLattice {}

impl HipLattice {
    /// `Lattice::new` (src/momentum/lattice.rs:27-207): validates the case,
    /// builds the node classes on the device and uploads the initial state.
    pub fn new(spec: &CaseSpec) -> LbResult<Self> {
        let (d, q) = match spec.velocity_set {
            sys::LBF_D2Q9 => (2, 9),
            sys::LBF_D3Q15 => (3, 15),
            sys::LBF_D3Q19 => (3, 19),
            sys::LBF_D3Q27 => (3, 27),
            v => return Err(format!("velocity set {v} is not constructible").into()),
        };
        let n_nodes = spec.n[..d].iter().map(|&v| v as usize).product();
        let mut faces = [sys::lbf_face_bc::default(); 6];
        for (f, bc) in faces.iter_mut().zip(spec.faces.iter()).take(2 * d) {
            *f = bc.to_c();
        }
        let (force_kind, force, force_field) = match (&spec.force_field, spec.force) {
            (Some(field), _) => {
                if field.len() != n_nodes * d {
                    return Err("force field needs nodes x d values".into());
                }
                (sys::LBF_SOURCE_FIELD, [0.0; 3], field.as_ptr())
            }
            (None, Some(f)) => (sys::LBF_SOURCE_UNIFORM, f, ptr::null()),
            (None, None) => (sys::LBF_SOURCE_NONE, [0.0; 3], ptr::null()),
        };
        if let Some(t) = &spec.node_type {
            if t.len() != n_nodes {
                return Err("node_type needs one entry per node".into());
            }
        }
        let desc = sys::lbf_case_desc {
            abi_version: sys::LBF_ABI_VERSION,
            velocity_set: spec.velocity_set,
            n: spec.n,
            flags: spec.flags,
            collision: spec.collision.to_c(),
            node_type: spec.node_type.as_ref().map_or(ptr::null(), |t| t.as_ptr()),
            faces,
            force_kind,
            precision: spec.precision,
            force,
            force_field,
            initial_density: ptr::null(),
            initial_density_uniform: spec.initial_density,
            initial_velocity: ptr::null(),
            initial_velocity_uniform: spec.initial_velocity,
            n_scalars: 0,
            streaming: spec.streaming,
            scalars: ptr::null(),
            shift_scale: 0.0,
            multiphase: 0,
            reserved: 0,
            g_function: 0.0,
            wall_density: 0.0,
        };
        let mut handle = ptr::null_mut();
        // SAFETY: desc and the arrays it points to outlive the call; the
        // library copies what it keeps.
        check(unsafe { sys::lbf_create(&desc, &mut handle) })?;
        Ok(HipLattice { handle, d, n_nodes, q })
    }

    /// `Lattice::initialize_nodes` (src/momentum/lattice.rs:322-328).
    pub fn initialize_nodes(&mut self) -> LbResult<()> {
        check(unsafe { sys::lbf_initialize_nodes(self.handle) })
    }

    /// `nsteps` x `Lattice::main_steps` (src/momentum/lattice.rs:431-441),
    /// with residual norms after every step when `residuals` is set.
    pub fn main_steps(&mut self, nsteps: u32, residuals: bool) -> LbResult<()> {
        let flags = if residuals { sys::LBF_STEP_RESIDUALS } else { 0 };
        check(unsafe { sys::lbf_main_steps(self.handle, nsteps, flags) })
    }

    /// The time loop with its stop rule (src/momentum.rs:273-292): at most
    /// max_iterations + 2 steps. Returns the number of completed steps.
    pub fn run(&mut self, max_iterations: u64) -> LbResult<u64> {
        let mut steps = 0u64;
        check(unsafe { sys::lbf_run(self.handle, max_iterations, &mut steps) })?;
        Ok(steps)
    }

    /// `compute_lattice_residuals` of the last step (src/momentum/lattice.rs:369-394):
    /// (|d rho|_2, |d u_x|_2 .. per dimension).
    pub fn residuals(&mut self) -> LbResult<(f64, Vec<f64>)> {
        let mut r = 0.0;
        let mut ru = [0.0f64; 3];
        check(unsafe {
            sys::lbf_residuals(self.handle, &mut r, ru.as_mut_ptr(), ptr::null_mut())
        })?;
        Ok((r, ru[..self.d].to_vec()))
    }

    /// `Lattice::time_step`.
    pub fn time_step(&mut self) -> LbResult<u64> {
        let mut t = 0u64;
        check(unsafe { sys::lbf_time_step(self.handle, &mut t) })?;
        Ok(t)
    }

    /// `compute_mean_density` (src/momentum/post.rs:43-58), summed in FP64 on the device.
    pub fn mean_density(&mut self) -> LbResult<f64> {
        let mut v = 0.0;
        check(unsafe { sys::lbf_reduce(self.handle, sys::LBF_REDUCE_MEAN_DENSITY, &mut v) })?;
        Ok(v)
    }

    /// Density [nodes] and velocity [nodes x d] in the reference's node order.
    pub fn fields(&mut self) -> LbResult<(Vec<f64>, Vec<f64>)> {
        let mut rho = vec![0.0; self.n_nodes];
        let mut u = vec![0.0; self.n_nodes * self.d];
        check(unsafe { sys::lbf_get_fields(self.handle, rho.as_mut_ptr(), u.as_mut_ptr()) })?;
        Ok((rho, u))
    }

    /// Post-stream populations [nodes x q] (the nodes' `f` vectors).
    pub fn populations(&mut self) -> LbResult<Vec<f64>> {
        let mut f = vec![0.0; self.n_nodes * self.q];
        check(unsafe { sys::lbf_get_pops(self.handle, 0, f.as_mut_ptr()) })?;
        Ok(f)
    }
}

impl Drop for HipLattice {
    fn drop(&mut self) {
        // SAFETY: the handle came from lbf_create and is destroyed once.
        unsafe { sys::lbf_destroy(self.handle) };
    }
}
