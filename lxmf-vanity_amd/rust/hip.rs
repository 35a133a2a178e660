//! The MI355X (gfx950) backend: the kernels are HIP, compiled ahead of time
//! into liblxmf_vanity_hip.so; this is the trait over its C ABI.
use super::{Backend, LaunchArgs};
use std::ffi::CStr;
use std::os::raw::{c_char, c_int};

// The first fields of lxv_open_opts; the library accepts this shorter struct and
// leaves `subwalks` and `walk` at their defaults (one walk per lane, the pair walk).
#[repr(C)]
struct OpenOpts { size: u32, device: i32, threads: u32, batch: u32, waves_per_cu: u32, layout: u32 }

#[repr(C)]
struct CLaunchArgs {
    iters: u32, ed: [u32; 8], name_hash: [u32; 3], want: [u32; 4], mask: [u32; 4],
    q: [u32; 8], max_hits: u32, dump_threads: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
struct Hit { thread: u32, reserved: u32, n: u64, address: [u32; 4] }

enum Ctx {}

extern "C" {
    fn lxv_open(opts: *const OpenOpts, out: *mut *mut Ctx) -> c_int;
    fn lxv_describe(ctx: *mut Ctx) -> *const c_char;
    fn lxv_threads(ctx: *mut Ctx) -> u32;
    fn lxv_batch(ctx: *mut Ctx) -> u64;
    fn lxv_init(ctx: *mut Ctx, u_prev_cur: *const u32, n_words: usize) -> c_int;
    fn lxv_launch(ctx: *mut Ctx, args: *const CLaunchArgs) -> c_int;
    fn lxv_poll(ctx: *mut Ctx, wait: c_int, hits: *mut Hit, cap: u32, n_hits: *mut u32,
                dump: *mut u32, dump_words: usize) -> c_int;
    fn lxv_close(ctx: *mut Ctx);
    fn lxv_last_error(ctx: *mut Ctx) -> *const c_char;
}

pub struct Hip { ctx: *mut Ctx }

fn err(ctx: *mut Ctx) -> String {
    unsafe { CStr::from_ptr(lxv_last_error(ctx)) }.to_string_lossy().into_owned()
}

impl Hip {
    pub fn new() -> Result<Hip, String> {
        // batch and waves_per_cu 0: the autotuner's single-search pick compiled
        // into the library (B = 256, 16 waves per CU, 8x32 limbs on MI355X).
        // threads: the reference's init computes every lane's start on the CPU
        // (first_public_keys), so 2^21 lanes (8 groups per resident wave);
        // 0 gives 2^24, for callers that set up lanes with lxv_init_seeded.
        let o = OpenOpts { size: std::mem::size_of::<OpenOpts>() as u32, device: 0, threads: 1 << 21,
                           batch: 0, waves_per_cu: 0, layout: 0 };
        let mut ctx = std::ptr::null_mut();
        if unsafe { lxv_open(&o, &mut ctx) } != 0 {
            return Err(err(std::ptr::null_mut()));
        }
        Ok(Hip { ctx })
    }
}

impl Drop for Hip {
    fn drop(&mut self) { unsafe { lxv_close(self.ctx) } }
}

impl Backend for Hip {
    fn describe(&self) -> String {
        unsafe { CStr::from_ptr(lxv_describe(self.ctx)) }.to_string_lossy().into_owned()
    }
    fn threads(&self) -> u32 { unsafe { lxv_threads(self.ctx) } }
    fn batch(&self) -> u64 { unsafe { lxv_batch(self.ctx) } }
    fn init(&mut self, u_prev_cur: &[u32]) -> Result<(), String> {
        match unsafe { lxv_init(self.ctx, u_prev_cur.as_ptr(), u_prev_cur.len()) } {
            0 => Ok(()),
            _ => Err(err(self.ctx)),
        }
    }
    fn launch(&mut self, a: &LaunchArgs) -> Result<(Vec<(u32, u64)>, Vec<u32>), String> {
        let c = CLaunchArgs { iters: a.iters, ed: a.ed, name_hash: a.name_hash, want: a.want,
                              mask: a.mask, q: a.q, max_hits: a.max_hits, dump_threads: a.dump_threads };
        if unsafe { lxv_launch(self.ctx, &c) } != 0 {
            return Err(err(self.ctx));
        }
        let mut hits = vec![Hit::default(); a.max_hits as usize];
        let mut dump = vec![0u32; a.dump_threads as usize * self.batch() as usize * 8];
        let mut n = 0u32;
        let rc = unsafe {
            lxv_poll(self.ctx, 1, hits.as_mut_ptr(), a.max_hits, &mut n,
                     if dump.is_empty() { std::ptr::null_mut() } else { dump.as_mut_ptr() }, dump.len())
        };
        if rc != 0 {
            return Err(err(self.ctx));
        }
        hits.truncate(n.min(a.max_hits) as usize);
        Ok((hits.iter().map(|h| (h.thread, h.n)).collect(), dump))
    }
}
