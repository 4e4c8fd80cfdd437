// The three insertions in gzippy's src/decompress/mod.rs that route single-
// member input to the GPU (the recipe of its DecodePath comment, mod.rs:50-53:
// a variant, a condition in classify_gzip, a dispatch arm).  Not compiled in
// this repository.  Cargo.toml gains
//     gzgpu-sys = { path = "<gzippy_amd>/integration/gzgpu-sys", optional = true }
//     [features] gpu = ["dep:gzgpu-sys"]
// and src/decompress/parallel/mod.rs gains `#[cfg(feature = "gpu")] pub mod gpu_sm;`.

// (1) in `pub enum DecodePath` (mod.rs:61):
    #[cfg(feature = "gpu")]
    GpuParallelSM,

// (2) in `classify_gzip` (mod.rs:158), where the ParallelSM result is
//     returned (after the multi-member and StoredParallel conditions):
    #[cfg(feature = "gpu")]
    if crate::decompress::parallel::gpu_sm::gpu_available() {
        return DecodePath::GpuParallelSM;
    }

// (3) in `decompress_single_member_for<W: Write>` (mod.rs:629), beside the
//     ParallelSM arm (mod.rs:634-650):
        #[cfg(feature = "gpu")]
        DecodePath::GpuParallelSM => {
            let n = crate::decompress::parallel::gpu_sm::decompress_parallel_gpu(data, writer)
                .map_err(|e| GzippyError::decompression(format!("parallel SM (gpu): {e}")))?;
            writer.flush()?;
            Ok(n)
        }
