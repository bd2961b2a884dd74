// ragged batches: see nfft_ragged.hip (added with the persistent kernel)
