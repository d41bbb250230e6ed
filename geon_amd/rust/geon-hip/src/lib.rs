//! `geon-hip`: the MI355X backend of geon's geodesic path as a drop-in.
//!
//! [`HipMetric`] wraps one of geon's four metrics and implements
//! `geon::metric::Metric` by delegating to it, so it goes wherever a
//! `Metric` goes: `Camera::tetrad`, `geodesic::trace`, the `curvature`
//! functions and `render` of the reference run on the CPU unchanged. The
//! functions of this crate with the reference's names and signatures —
//! [`render`], [`trace`], [`riemann`], [`ricci`], [`ricci_scalar`],
//! [`einstein`], [`stress_energy`], [`nec_contraction`] — run the same
//! computation on the GPU through libgeon_hip (one process-wide context on
//! device 0, calls serialised). Failures the reference reports by panicking
//! (camera.rs:53,85; types.rs:39) panic here too, with the library's message.
//!
//! ```ignore
//! use geon::{Camera, RenderConfig};
//! use geon::metric::schwarzschild::Schwarzschild;
//! use geon_hip::HipMetric;
//! let m = HipMetric::Schwarzschild(Schwarzschild::new(1.0));
//! let cam = Camera::new([0.0, 20.0, std::f64::consts::FRAC_PI_2, 0.0], 60.0, 1024, 1024);
//! let (pixels, stats) = geon_hip::render(&m, &cam, &RenderConfig::default());
//! ```

use std::ffi::CStr;
use std::ptr;
use std::sync::atomic::Ordering;
use std::sync::{Mutex, OnceLock};

use geon::camera::Camera;
use geon::curvature::Riemann;
use geon::geodesic::TraceResult;
use geon::integrator::State;
use geon::metric::alcubierre::Alcubierre;
use geon::metric::minkowski::Minkowski;
use geon::metric::morris_thorne::MorrisThorne;
use geon::metric::schwarzschild::Schwarzschild;
use geon::metric::{Metric, Termination};
use geon::render::{RenderConfig, RenderStats};
use geon::types::{Christoffel, Mat4, Scalar, Vec4};
use geon_hip_sys as sys;

/// A metric this library runs on the GPU.
pub enum HipMetric {
    Minkowski(Minkowski),
    Schwarzschild(Schwarzschild),
    Alcubierre(Alcubierre),
    MorrisThorne(MorrisThorne),
}

impl HipMetric {
    fn inner(&self) -> &dyn Metric {
        match self {
            HipMetric::Minkowski(m) => m,
            HipMetric::Schwarzschild(m) => m,
            HipMetric::Alcubierre(m) => m,
            HipMetric::MorrisThorne(m) => m,
        }
    }

    /// The ABI description: kind, parameters and the termination radii the
    /// reference's `terminate` uses (r² > 1e6 and |l| > 1000 are escape
    /// radius 1000 in the ABI's terms).
    pub fn to_c(&self) -> sys::geon_metric_t {
        let mut m = sys::geon_metric_t::default();
        let kind = match self {
            HipMetric::Minkowski(_) => sys::GEON_METRIC_MINKOWSKI,
            HipMetric::Schwarzschild(_) => sys::GEON_METRIC_SCHWARZSCHILD,
            HipMetric::Alcubierre(_) => sys::GEON_METRIC_ALCUBIERRE,
            HipMetric::MorrisThorne(_) => sys::GEON_METRIC_MORRIS_THORNE,
        };
        check(unsafe { sys::geon_metric_default(kind, &mut m) }, "geon_metric_default");
        match self {
            HipMetric::Minkowski(_) => {}
            HipMetric::Schwarzschild(s) => {
                m.p[0] = s.mass;
                m.horizon_pad = s.horizon_pad;
                m.escape_radius = s.escape_radius;
            }
            HipMetric::Alcubierre(a) => {
                m.p[0] = a.bubble_radius;
                m.p[1] = a.bubble_thickness;
                m.p[2] = a.ship_velocity;
                m.p[3] = a.ship_x;
            }
            HipMetric::MorrisThorne(w) => m.p[0] = w.throat_radius,
        }
        m.fd_step = self.inner().fd_step(&[0.0; 4])[0];
        m
    }
}

impl Metric for HipMetric {
    fn g(&self, x: &Vec4) -> Mat4 {
        self.inner().g(x)
    }
    fn dg(&self, x: &Vec4) -> [Mat4; 4] {
        self.inner().dg(x)
    }
    fn g_inv(&self, x: &Vec4) -> Mat4 {
        self.inner().g_inv(x)
    }
    fn christoffel(&self, x: &Vec4) -> Christoffel {
        self.inner().christoffel(x)
    }
    fn fd_step(&self, x: &Vec4) -> [Scalar; 4] {
        self.inner().fd_step(x)
    }
    fn terminate(&self, x: &Vec4) -> Option<Termination> {
        self.inner().terminate(x)
    }
    fn name(&self) -> &'static str {
        self.inner().name()
    }
    fn asymptotic_direction(&self, x: &Vec4, p: &Vec4) -> [Scalar; 3] {
        self.inner().asymptotic_direction(x, p)
    }
}

fn check(status: i32, what: &str) {
    if status != sys::GEON_OK {
        let msg = unsafe { CStr::from_ptr(sys::geon_strerror(status)) }.to_string_lossy();
        let detail = unsafe { CStr::from_ptr(sys::geon_last_error()) }.to_string_lossy();
        panic!("{}: {} ({}) {}", what, msg, status, detail);
    }
}

struct Ctx(*mut sys::geon_ctx_t);
// The context is used by one thread at a time (the mutex below).
unsafe impl Send for Ctx {}

fn with_ctx<R>(f: impl FnOnce(*mut sys::geon_ctx_t) -> R) -> R {
    static CTX: OnceLock<Mutex<Ctx>> = OnceLock::new();
    let m = CTX.get_or_init(|| {
        let mut h: *mut sys::geon_ctx_t = ptr::null_mut();
        check(unsafe { sys::geon_ctx_create(0, ptr::null(), &mut h) }, "geon_ctx_create");
        Mutex::new(Ctx(h))
    });
    let guard = m.lock().unwrap_or_else(|e| e.into_inner());
    f(guard.0)
}

fn camera_c(camera: &Camera) -> sys::geon_camera_t {
    sys::geon_camera_t {
        observer: camera.observer,
        fov_y_rad: camera.fov_y_rad,
        width: u32::try_from(camera.width).expect("width fits u32"),
        height: u32::try_from(camera.height).expect("height fits u32"),
    }
}

fn termination(cls: i32) -> Termination {
    match cls {
        sys::GEON_CAPTURED => Termination::Captured,
        sys::GEON_ESCAPED => Termination::Escaped,
        _ => Termination::Diverged, // diverged, or out of steps (geodesic.rs:107-116)
    }
}

/// `render::render` (render.rs:55-165) on the GPU: linear RGB, row-major,
/// top-left origin, and the frame statistics; `sky::SHOW_STARS` is honoured.
pub fn render(metric: &HipMetric, camera: &Camera, cfg: &RenderConfig) -> (Vec<f32>, RenderStats) {
    let m = metric.to_c();
    let cam = camera_c(camera);
    let c = sys::geon_render_cfg_t {
        step: cfg.step,
        max_steps: u32::try_from(cfg.max_steps).expect("max_steps fits u32"),
        samples_per_axis: u32::try_from(cfg.samples_per_axis.max(1)).expect("samples fit u32"),
        captured_rgb: cfg.captured_rgb,
        diverged_rgb: cfg.diverged_rgb,
        show_stars: geon::sky::SHOW_STARS.load(Ordering::Relaxed) as u32,
        precision: sys::GEON_PREC_FP64,
        fp64_band: 0.0,
    };
    let mut pixels = vec![0.0f32; 3 * camera.width * camera.height];
    let mut st = sys::geon_render_stats_t::default();
    with_ctx(|ctx| {
        check(unsafe {
                  sys::geon_render(ctx, &m, &cam, &c, pixels.as_mut_ptr(), ptr::null_mut(),
                                   ptr::null_mut(), &mut st)
              },
              "render")
    });
    let stats = RenderStats {
        captured: st.captured as usize,
        escaped: st.escaped as usize,
        diverged: st.diverged as usize,
        max_null_residual: st.max_null_residual,
        avg_steps: st.avg_steps,
    };
    (pixels, stats)
}

/// `geodesic::trace` (geodesic.rs:62-117) for many rays at once on the GPU.
pub fn trace_many(metric: &HipMetric, y0: &[State], step: Scalar, max_steps: usize)
    -> Vec<TraceResult> {
    let n = y0.len();
    if n == 0 {
        return Vec::new();
    }
    let mut soa = vec![0.0f64; 8 * n];
    for (k, y) in y0.iter().enumerate() {
        for c in 0..8 {
            soa[c * n + k] = y[c];
        }
    }
    let (mut cls, mut steps) = (vec![0i32; n], vec![0u32; n]);
    let (mut yf, mut resid) = (vec![0.0f64; 8 * n], vec![0.0f64; n]);
    let m = metric.to_c();
    let max = u32::try_from(max_steps).expect("max_steps fits u32");
    with_ctx(|ctx| {
        check(unsafe {
                  sys::geon_trace_rays(ctx, &m, soa.as_ptr(), n, step, max, cls.as_mut_ptr(),
                                       steps.as_mut_ptr(), yf.as_mut_ptr(), resid.as_mut_ptr(),
                                       ptr::null_mut())
              },
              "trace")
    });
    (0..n)
        .map(|k| {
            let mut s = [0.0f64; 8];
            for c in 0..8 {
                s[c] = yf[c * n + k];
            }
            TraceResult {
                termination: termination(cls[k]),
                final_state: s,
                steps_taken: steps[k] as usize,
                null_residual: resid[k],
            }
        })
        .collect()
}

/// `geodesic::trace` (geodesic.rs:62-117) of one ray on the GPU.
pub fn trace(metric: &HipMetric, y0: State, step: Scalar, max_steps: usize) -> TraceResult {
    trace_many(metric, &[y0], step, max_steps).pop().expect("one result")
}

fn curvature(metric: &HipMetric, x: &Vec4, what: i32, width: usize) -> Vec<f64> {
    let m = metric.to_c();
    let mut out = vec![0.0f64; width];
    // one point: its SoA layout is the point itself
    with_ctx(|ctx| {
        check(unsafe { sys::geon_curvature(ctx, &m, x.as_ptr(), 1, what, out.as_mut_ptr()) },
              "curvature")
    });
    out
}

fn mat4(v: &[f64]) -> Mat4 {
    let mut m = [[0.0; 4]; 4];
    for i in 0..4 {
        for j in 0..4 {
            m[i][j] = v[4 * i + j];
        }
    }
    m
}

/// `curvature::riemann` (curvature.rs:48-70) on the GPU.
pub fn riemann(metric: &HipMetric, x: &Vec4) -> Riemann {
    let v = curvature(metric, x, sys::GEON_CURV_RIEMANN, 256);
    let mut r = [[[[0.0; 4]; 4]; 4]; 4];
    for a in 0..4 {
        for b in 0..4 {
            for c in 0..4 {
                for d in 0..4 {
                    r[a][b][c][d] = v[64 * a + 16 * b + 4 * c + d];
                }
            }
        }
    }
    r
}

/// `curvature::ricci` (curvature.rs:72-85) on the GPU.
pub fn ricci(metric: &HipMetric, x: &Vec4) -> Mat4 {
    mat4(&curvature(metric, x, sys::GEON_CURV_RICCI, 16))
}

/// `curvature::ricci_scalar` (curvature.rs:87-97) on the GPU.
pub fn ricci_scalar(metric: &HipMetric, x: &Vec4) -> Scalar {
    curvature(metric, x, sys::GEON_CURV_SCALAR, 1)[0]
}

/// `curvature::einstein` (curvature.rs:99-110) on the GPU.
pub fn einstein(metric: &HipMetric, x: &Vec4) -> Mat4 {
    mat4(&curvature(metric, x, sys::GEON_CURV_EINSTEIN, 16))
}

/// `curvature::stress_energy` (curvature.rs:112-125) on the GPU.
pub fn stress_energy(metric: &HipMetric, x: &Vec4) -> Mat4 {
    mat4(&curvature(metric, x, sys::GEON_CURV_STRESS, 16))
}

/// `curvature::nec_contraction` (curvature.rs:127-136) on the GPU.
pub fn nec_contraction(metric: &HipMetric, x: &Vec4, k: &Vec4) -> Scalar {
    let m = metric.to_c();
    let mut out = 0.0f64;
    with_ctx(|ctx| {
        check(unsafe { sys::geon_nec_contraction(ctx, &m, x.as_ptr(), 1, k.as_ptr(), &mut out) },
              "nec_contraction")
    });
    out
}

/// `curvature::wec_contraction` (curvature.rs:139-141) on the GPU: the same
/// contraction with the observer's 4-velocity.
pub fn wec_contraction(metric: &HipMetric, x: &Vec4, u: &Vec4) -> Scalar {
    nec_contraction(metric, x, u)
}
