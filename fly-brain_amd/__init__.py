"""fly-brain_amd: MI355X-native engine for the fly-brain connectome LIF step and its adjoint."""
