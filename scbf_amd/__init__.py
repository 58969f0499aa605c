"""MI355X-native SuperCountingBloom engine: hand-written gfx950 HIP kernels
behind a C ABI (include/scb.h, scbf_amd/libscb.so), with this Python mirror of
the super-counting-bloom crate API on top."""
from .api import (METHOD_AUTO, METHOD_BATCHED, METHOD_PLAIN, REQ_ABUNDANCES, REQ_SPECTRUM, AddReport,
                  ApproxSpectrum, BuildIndexReport, CounterBits, FilterOnlySpectrumEstimate,
                  QueryFastaReport, SampledSpectrum, SequenceAbundanceSummary, ServeConfig, Served, SpectrumAccumulator,
                  StreamEstimateReport, StreamEstimateStats, SuperCountingBloomBuilder,
                  SuperCountingBloomConfig, SuperCountingBloomError, SuperCountingBloomIndex,
                  SuperCountingBloomReport, SuperCountingBloomServer, SuperCountingBloomStats, Tuning,
                  canonical_kmer, concat_records, device_count, estimate_spectrum_from_fastx,
                  estimate_spectrum_from_fastx_pair, filter_only_count_hint_from_histogram,
                  filter_only_spectrum_from_histogram, for_each_fastx_record, host_array,
                  sampled_spectrum_from_fastx, tuning_cache_load,
                  summarize_abundances,
                  stream_estimates_from_fastx, stream_estimates_from_fastx_pair, write_spectrum,
                  xorshift_u64)
