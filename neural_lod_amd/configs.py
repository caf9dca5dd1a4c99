"""Model configurations of the neural-LoD networks (tcnn JSON as Python dicts).

``VISIBILITY`` / ``THROUGHPUT`` are the benchmark workloads (hash grid 16x2
features, 2^19 entries per level, base 16, growth 1.5; SH degree 4; MLP width
64).  ``REFERENCE_*`` restate the configurations neural_lod builds in
rptr/cuda/neural-lod/neural_lod_learning_init.cu:577-676,822-911 with the
scene settings of neural_lod/scenes/arcade/arcade_world_final.ini:141-254.
"""

HASH_GRID = {
    "otype": "HashGrid",
    "n_levels": 16,
    "n_features_per_level": 2,
    "log2_hashmap_size": 19,
    "base_resolution": 16,
    "per_level_scale": 1.5,
    "interpolation": "Linear",
}

# inputs: position (3) + direction mapped to [0,1]^3 (3)
VISIBILITY = {
    "n_input_dims": 6,
    "n_output_dims": 1,
    "encoding": {
        "otype": "Composite",
        "nested": [
            dict(HASH_GRID, n_dims_to_encode=3),
            {"otype": "SphericalHarmonics", "n_dims_to_encode": 3, "degree": 4},
        ],
    },
    "network": {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "Sigmoid",
                "n_neurons": 64, "n_hidden_layers": 4},
    "loss": {"otype": "BinaryCrossEntropy"},
    "optimizer": {"otype": "Adam", "learning_rate": 1e-2, "beta1": 0.9, "beta2": 0.99, "epsilon": 1e-15,
                  "l2_reg": 1e-8},
}

# inputs: position (3), outgoing direction (3), incident direction (3), LoD (1)
THROUGHPUT = {
    "n_input_dims": 10,
    "n_output_dims": 3,
    "encoding": {
        "otype": "Composite",
        "nested": [
            dict(HASH_GRID, n_dims_to_encode=3),
            {"otype": "SphericalHarmonics", "n_dims_to_encode": 3, "degree": 4},
            {"otype": "SphericalHarmonics", "n_dims_to_encode": 3, "degree": 4},
            {"otype": "Identity", "n_dims_to_encode": 1},
        ],
    },
    "network": {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "Exponential",
                "n_neurons": 64, "n_hidden_layers": 5},
    "loss": {"otype": "RelativeL2"},
    "optimizer": {"otype": "Adam", "learning_rate": 1e-2, "beta1": 0.9, "beta2": 0.99, "epsilon": 1e-15,
                  "l2_reg": 1e-8},
}

_REF_GRID = {"otype": "HashGrid", "n_levels": 7, "n_features_per_level": 2, "log2_hashmap_size": 19,
             "base_resolution": 8, "per_level_scale": 2.0, "interpolation": "Linear"}

REFERENCE_VISIBILITY = {
    "n_input_dims": 6,
    "n_output_dims": 1,
    "encoding": {"otype": "Composite", "n_dims_to_encode": 6, "nested": [
        {"otype": "RepeatedComposite", "n_repetitions": 2, "n_dims_to_encode": 6,
         "nested": [dict(_REF_GRID, n_dims_to_encode=3)]}]},
    "network": {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "Sigmoid",
                "n_neurons": 128, "n_hidden_layers": 4},
    "loss": {"otype": "BinaryCrossEntropy"},
    "optimizer": {"otype": "ExponentialDecay", "decay_start": 8000, "decay_interval": 4000, "decay_base": 0.33,
                  "nested": {"otype": "Adam", "learning_rate": 5e-3, "beta1": 0.9, "beta2": 0.99,
                             "epsilon": 1e-15, "l2_reg": 1e-8}},
}

REFERENCE_THROUGHPUT = {
    "n_input_dims": 9,
    "n_output_dims": 3,
    "encoding": {"otype": "Composite", "nested": [
        dict(_REF_GRID, n_dims_to_encode=3),
        {"otype": "SphericalHarmonics", "n_dims_to_encode": 3, "degree": 8},
        {"otype": "SphericalHarmonics", "n_dims_to_encode": 3, "degree": 8}]},
    "network": {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "None",
                "n_neurons": 128, "n_hidden_layers": 4},
    "loss": {"otype": "RelativeL2"},
    "optimizer": {"otype": "Ema", "decay": 0.95, "nested": {
        "otype": "ExponentialDecay", "decay_start": 8000, "decay_interval": 4000, "decay_base": 0.33,
        "nested": {"otype": "Adam", "learning_rate": 5e-3, "beta1": 0.9, "beta2": 0.99, "epsilon": 1e-15,
                   "l2_reg": 1e-8}}},
}

ALL = {"visibility": VISIBILITY, "throughput": THROUGHPUT,
       "reference_visibility": REFERENCE_VISIBILITY, "reference_throughput": REFERENCE_THROUGHPUT}
