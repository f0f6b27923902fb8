"""opencloth_amd -- MI355X-native implicit mass-spring cloth step."""
