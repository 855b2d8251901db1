! placeholder
