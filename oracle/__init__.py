"""Oracle package: CPU restatement of ChannelFlow.jl's step path (test infrastructure only)."""
